// Device-side building blocks shared by the gfx950 kernels: the POD
// configuration, the vector-field registry with Jacobians and parameter
// Jacobians, Taylor-coefficient initialisation, and wavefront-group
// primitives (a group is G consecutive lanes of one wavefront that own one
// IVP instance).
#pragma once

#ifdef PDQ_HOST_EMU
#ifndef PDQ_EMU_SHIM_INCLUDED
#include "pdq_hostemu.h"
#endif
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "../../include/pdq.h"

namespace pdq {

constexpr int kMaxParams = 64;   // vector-field parameters staged in LDS
constexpr int kMaxTangents = 8;  // forward-mode directions of the LML gradient
// Safety net that bounds every instance's loop (the reference has none).
constexpr long kMaxTotalAttempts = 1L << 22;  // per solve: every persistent wave ends in bounded time

// ------------------------------------------------- refined reciprocal and root
// Reciprocal, quotient and square root of operands known to be finite and
// non-zero, from the hardware estimates refined by FMA iterations.  The IEEE
// sequences the compiler emits for 1/b and sqrt(x) are 9-15 instructions with
// long dependency chains; the Householder reflectors and the scalings of the
// hot loops need dozens of them per step.
// * lane_rcp_nz: v_rcp_f64, then two Newton steps y += y (1 - b y), each an
//   FMA pair; within an ulp of 1/b for finite non-zero b (zero or infinite b
//   give NaN: callers pass such y only to a checked division).
// * lane_sqrt_pos: v_rsq_f64, one Goldschmidt step for (sqrt x, 1/(2 sqrt x))
//   and a final Newton correction of the root (Markstein, IA-64 and
//   Elementary Functions, 2000, ch. 9); within an ulp of sqrt(x) for finite
//   positive x.
// * lane_div_nz: q = a / b from y = lane_rcp_nz(b): q = a y, corrected by one
//   FMA step against the residual a - b q (Markstein); within an ulp for
//   finite non-zero b and a finite quotient.
// The host build (tests/emu) computes the IEEE operations.
#ifdef PDQ_HOST_EMU
__device__ __forceinline__ double lane_rcp_nz(double b) { return 1.0 / b; }
__device__ __forceinline__ double lane_sqrt_pos(double x) { return sqrt(x); }
#else
__device__ __forceinline__ double lane_rcp_nz(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double lane_sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  return fma(fma(-g, g, x), h, g);
}
#endif
__device__ __forceinline__ double lane_div_nz(double a, double b, double y) {
  const double q = a * y;
  return fma(fma(-q, b, a), y, q);
}

// The inverse N-th root, from a hardware estimate refined by FMA iterations.
// * lane_root_inv<N>: x^(-1/N) from a single-precision estimate of the
//   scaled operand (x = xs 2^(N k), xs within single range), then two Newton
//   steps on x y^N = 1, y += (y / N) (1 - x y^N); within a few ulps of
//   pow(x, -1/N).  x = 0 gives +inf and x = +inf gives 0, as pow does.
// The host build (tests/emu) computes pow.
#ifdef PDQ_HOST_EMU
template <int N>
__device__ __forceinline__ double lane_root_inv(double x) {
  return pow(x, -1.0 / (double)N);
}
#else
template <int N>
__device__ __forceinline__ double lane_root_inv(double x) {
  const int k = __builtin_amdgcn_frexp_exp(x) / N;
  const double xs = __builtin_ldexp(x, -N * k);
  double y = __builtin_ldexp((double)__builtin_amdgcn_exp2f(-__builtin_amdgcn_logf((float)xs) / (float)N), -k);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    double t = y;
#pragma unroll
    for (int q = 1; q < N; ++q) t *= y;
    y = fma(y * (1.0 / (double)N), fma(-x, t, 1.0), y);
  }
  y = x == 0.0 ? __builtin_inf() : y;
  return x == __builtin_inf() ? 0.0 : y;
}
#endif

// preconditioner_taylor (utilities.py:74-84): dt^k with libm pow semantics.
// dt^k for an exponent known at compile time, computed as an unevaluated sum
// hi + lo of exact FMA products (double-double), then rounded once: the
// result is within about half an ulp of dt^k, as a correctly rounded pow is,
// at a few instructions instead of a general pow.  k = 0, 1, 2, -1 are single
// correctly rounded operations.  A negative exponent is the reciprocal of the
// double-double power: q = 1/hi refined by one FMA step against hi + lo.
struct LaneDD {
  double hi, lo;
};

template <int k>
__device__ __forceinline__ LaneDD lane_dd_pow(double x) {
  static_assert(k >= 1 && k <= 8, "lane_dd_pow covers the exponents of num_derivatives <= 7");
  if constexpr (k == 1) {
    return {x, 0.0};
  } else if constexpr (k % 2 == 0) {
    const LaneDD h = lane_dd_pow<k / 2>(x);
    const double p = h.hi * h.hi;
    return {p, fma(h.hi, h.hi, -p) + 2.0 * h.hi * h.lo};
  } else {
    const LaneDD h = lane_dd_pow<k - 1>(x);
    const double p = h.hi * x;
    return {p, fma(h.hi, x, -p) + h.lo * x};
  }
}

template <int k>
__device__ __forceinline__ double lane_ipow(double dt) {
  if constexpr (k == 0) {
    return 1.0;
  } else if constexpr (k == 1) {
    return dt;
  } else if constexpr (k == 2) {
    return dt * dt;
  } else if constexpr (k == -1) {
    return 1.0 / dt;
  } else if constexpr (k > 0) {
    const LaneDD p = lane_dd_pow<k>(dt);
    return p.hi + p.lo;
  } else {
    const LaneDD p = lane_dd_pow<-k>(dt);
    const double q = 1.0 / p.hi;
    const double r = fma(-p.hi, q, 1.0) - p.lo * q;
    return isfinite(q) ? fma(q, r, q) : q;  // dt^k = 0 or overflowed: 1/hi as pow gives it
  }
}

// ------------------------------------------------------------ vector fields
// The registry restates problems.py:283-295's u' = f(u, t) for the device.
// Component-wise evaluation lets every lane of a group produce its own f_k.
__device__ __forceinline__ double vf_component(int id, int d, const double* __restrict__ y,
                                               const double* __restrict__ prm, double t, int k) {
  switch (id) {
    case PDQ_VF_LOTKA_VOLTERRA:
      return k == 0 ? prm[0] * y[0] - prm[1] * y[0] * y[1] : -prm[2] * y[1] + prm[3] * y[0] * y[1];
    case PDQ_VF_VAN_DER_POL:
      return k == 0 ? y[1] : prm[0] * ((1.0 - y[0] * y[0]) * y[1]) - y[0];
    case PDQ_VF_BURGERS: {
      const double visc = prm[0], dx = prm[1];
      const double um = y[(k + d - 1) % d], u0 = y[k], up = y[(k + 1) % d];
      return -u0 * ((up - um) / (2.0 * dx)) + visc * ((up - 2.0 * u0 + um) / (dx * dx));
    }
    case PDQ_VF_LINEAR:
      return prm[k] * y[k] + t;
    case PDQ_VF_LOGISTIC:
      return prm[0] * y[k] * (1.0 - y[k]);
  }
  return 0.0;
}

// d f_k / d y_j
__device__ __forceinline__ double vf_jac(int id, int d, const double* __restrict__ y, const double* __restrict__ prm,
                                         int k, int j) {
  switch (id) {
    case PDQ_VF_LOTKA_VOLTERRA:
      if (k == 0) return j == 0 ? prm[0] - prm[1] * y[1] : -prm[1] * y[0];
      return j == 0 ? prm[3] * y[1] : -prm[2] + prm[3] * y[0];
    case PDQ_VF_VAN_DER_POL:
      if (k == 0) return j == 0 ? 0.0 : 1.0;
      return j == 0 ? prm[0] * (-2.0 * y[0] * y[1]) - 1.0 : prm[0] * (1.0 - y[0] * y[0]);
    case PDQ_VF_BURGERS: {
      const double visc = prm[0], dx = prm[1];
      const int km = (k + d - 1) % d, kp = (k + 1) % d;
      const double um = y[km], u0 = y[k], up = y[kp];
      double v = 0.0;
      if (j == k) v += -((up - um) / (2.0 * dx)) - 2.0 * visc / (dx * dx);
      if (j == kp) v += -u0 / (2.0 * dx) + visc / (dx * dx);
      if (j == km) v += u0 / (2.0 * dx) + visc / (dx * dx);
      return v;
    }
    case PDQ_VF_LINEAR:
      return j == k ? prm[k] : 0.0;
    case PDQ_VF_LOGISTIC:
      return j == k ? prm[0] * (1.0 - 2.0 * y[k]) : 0.0;
  }
  return 0.0;
}

// Rademacher probes of the Monte Carlo Jacobian handler (jacobian_monte_carlo_rev,
// jacobians.py:195-271): +1 or -1 from the top bit of the SplitMix64 finaliser
// (Steele, Lea and Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) applied to a counter that combines the seed, the probe set (the
// accepted steps before the linearisation), the probe and the coordinate.
__host__ __device__ inline double pdq_rademacher(uint32_t seed, uint32_t set, uint32_t probe, uint32_t coord) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)set * 0xD1B54A32D192ED03ull +
               (uint64_t)probe * 0x8CB92BA72F3D8DD7ull + (uint64_t)coord * 0xABC98388FB8FAC03ull;
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ull;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (z >> 63) ? -1.0 : 1.0;
}

// The probe set of a step's first linearisation after `steps` accepted steps.
// The reference carries the handler's PRNG key in the solution and splits it
// at every linearisation (jacobians.py:230-271): solver_dynamic with
// re_linearize_after_calibration linearises twice per step (solvers.py:
// 561-582), its second linearisation taking set + 1.  A rejected proposal's
// key is discarded with it, so a retry draws the same probes.
__host__ __device__ inline uint32_t pdq_lin_set(const pdq_config& cfg, long steps) {
  const uint32_t per = (cfg.solver == PDQ_SOLVER_DYNAMIC && cfg.re_linearize_after_calibration) ? 2u : 1u;
  return (uint32_t)steps * per;
}

// The rows j of the Jacobian's column k that can be non-zero (d f_j / d y_k),
// in ascending order: at most three.
__device__ __forceinline__ int vf_col_rows(int id, int d, int k, int* rows) {
  switch (id) {
    case PDQ_VF_LOTKA_VOLTERRA:
    case PDQ_VF_VAN_DER_POL:
      rows[0] = 0;
      rows[1] = 1;
      return 2;
    case PDQ_VF_BURGERS: {
      if (d <= 3) {
        for (int j = 0; j < d; ++j) rows[j] = j;
        return d;
      }
      const int km = (k + d - 1) % d, kp = (k + 1) % d;
      // ascending: the wrapped neighbour goes first (k = 0) or last (k = d - 1)
      if (k == 0) {
        rows[0] = 0, rows[1] = 1, rows[2] = km;
      } else if (k == d - 1) {
        rows[0] = kp, rows[1] = km, rows[2] = k;
      } else {
        rows[0] = km, rows[1] = k, rows[2] = kp;
      }
      return 3;
    }
  }
  rows[0] = k;
  return 1;
}

// The Monte Carlo estimate of the Jacobian's diagonal entry k: the mean over
// `probes` Rademacher vectors v of
//  * reverse mode (calculate_diagonal_along_d of jacobian_monte_carlo_rev,
//    jacobians.py:250-271): (v^T J)_k v_k = (sum_j v_j J[j][k]) v_k;
//  * forward mode (jacobian_monte_carlo_fwd, :171-192): (J v)_k v_k =
//    (sum_j J[k][j] v_j) v_k;
// with J[j][k] = d f_j / d y_k.  In both modes its sum over k is the trace
// estimate v^T J v of calculate_trace_along_d (:151-169, :230-248).  The
// registry's Jacobians have symmetric sparsity patterns, so row k and column k
// have their non-zeros at the same indices.
__device__ __forceinline__ double vf_mc_diag(int id, int d, const double* __restrict__ y,
                                             const double* __restrict__ prm, int k, uint32_t seed, uint32_t set,
                                             int probes, bool fwd) {
  int rows[3];
  const int nr = vf_col_rows(id, d, k, rows);
  double col[3];
  for (int r = 0; r < nr; ++r) col[r] = fwd ? vf_jac(id, d, y, prm, k, rows[r]) : vf_jac(id, d, y, prm, rows[r], k);
  double acc = 0.0;
  for (int s = 0; s < probes; ++s) {
    double w = 0.0;
    for (int r = 0; r < nr; ++r) w += pdq_rademacher(seed, set, (uint32_t)s, (uint32_t)rows[r]) * col[r];
    acc += w * pdq_rademacher(seed, set, (uint32_t)s, (uint32_t)k);
  }
  return acc / (double)probes;
}

// d f_k / d prm_q
__device__ __forceinline__ double vf_jac_param(int id, int d, const double* __restrict__ y,
                                               const double* __restrict__ prm, int k, int q) {
  switch (id) {
    case PDQ_VF_LOTKA_VOLTERRA:
      if (k == 0) return q == 0 ? y[0] : (q == 1 ? -(y[0] * y[1]) : 0.0);
      return q == 2 ? -y[1] : (q == 3 ? y[0] * y[1] : 0.0);
    case PDQ_VF_VAN_DER_POL:
      return (k == 1 && q == 0) ? (1.0 - y[0] * y[0]) * y[1] : 0.0;
    case PDQ_VF_BURGERS: {
      const double visc = prm[0], dx = prm[1];
      const double um = y[(k + d - 1) % d], u0 = y[k], up = y[(k + 1) % d];
      if (q == 0) return (up - 2.0 * u0 + um) / (dx * dx);
      return u0 * (up - um) / (2.0 * dx * dx) - 2.0 * visc * (up - 2.0 * u0 + um) / (dx * dx * dx);
    }
    case PDQ_VF_LINEAR:
      return q == k ? y[k] : 0.0;
    case PDQ_VF_LOGISTIC:
      return q == 0 ? y[k] * (1.0 - y[k]) : 0.0;
  }
  return 0.0;
}

// ----------------------------------------------------- Taylor coefficients
// Forward-mode numbers for the initial Taylor coefficients' sensitivities.
struct Dual {
  double v;
  double g[kMaxTangents];
};

__device__ __forceinline__ Dual dconst(double v) {
  Dual r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = 0.0;
  return r;
}
__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
__device__ __forceinline__ Dual operator/(const Dual& a, double s) {
  Dual r;
  r.v = a.v / s;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = a.g[i] / s;
  return r;
}
__device__ __forceinline__ Dual operator*(double s, const Dual& a) {
  Dual r;
  r.v = s * a.v;
#pragma unroll
  for (int i = 0; i < kMaxTangents; ++i) r.g[i] = s * a.g[i];
  return r;
}

__device__ __forceinline__ double tval(double x) { return x; }
__device__ __forceinline__ double tval(const Dual& x) { return x.v; }
__device__ __forceinline__ double tconst(double v, double) { return v; }
__device__ __forceinline__ Dual tconst(double v, const Dual&) { return dconst(v); }

// Normalised Taylor coefficient k of f(y(t0 + s)) given the coefficients
// c[0..k] (restates jet_expansion_algorithms.py:155-177 for the registry;
// products are Cauchy products).  c is (num+1) x d row-major.
template <class T>
__device__ T taylor_rhs(int id, int d, const T* c, const T* prm, double t0, int k, int j) {
  switch (id) {
    case PDQ_VF_LOTKA_VOLTERRA: {
      T s = tconst(0.0, c[0]);
      for (int i = 0; i <= k; ++i) s = s + c[i * d + 0] * c[(k - i) * d + 1];
      if (j == 0) return prm[0] * c[k * d + 0] - prm[1] * s;
      return (tconst(0.0, c[0]) - prm[2]) * c[k * d + 1] + prm[3] * s;
    }
    case PDQ_VF_VAN_DER_POL: {
      if (j == 0) return c[k * d + 1];
      T s = tconst(0.0, c[0]);
      for (int i = 0; i <= k; ++i) {
        T sq = tconst(0.0, c[0]);
        for (int a = 0; a <= i; ++a) sq = sq + c[a * d + 0] * c[(i - a) * d + 0];
        sq = tconst(i == 0 ? 1.0 : 0.0, c[0]) - sq;
        s = s + sq * c[(k - i) * d + 1];
      }
      return prm[0] * s - c[k * d + 0];
    }
    case PDQ_VF_BURGERS: {
      const int jm = (j + d - 1) % d, jp = (j + 1) % d;
      T s = tconst(0.0, c[0]);
      const double dx = tval(prm[1]);
      for (int i = 0; i <= k; ++i) s = s + c[i * d + j] * ((c[(k - i) * d + jp] - c[(k - i) * d + jm]) / (2.0 * dx));
      return prm[0] * ((c[k * d + jp] - 2.0 * c[k * d + j] + c[k * d + jm]) / (dx * dx)) - s;
    }
    case PDQ_VF_LINEAR:
      return prm[j] * c[k * d + j] + tconst(k == 0 ? t0 : (k == 1 ? 1.0 : 0.0), c[0]);
    case PDQ_VF_LOGISTIC: {
      T s = tconst(0.0, c[0]);
      for (int i = 0; i <= k; ++i) s = s + c[i * d + j] * c[(k - i) * d + j];
      return prm[0] * (c[k * d + j] - s);
    }
  }
  return tconst(0.0, c[0]);
}

__device__ __forceinline__ double factorial_d(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; ++i) f *= (double)i;
  return f;
}

// ------------------------------------------------------------- lane groups
template <int G>
struct Grp {
  int l;  // lane index inside the group
  // Orders LDS accesses of the group's lanes (they share one wavefront).
  __device__ __forceinline__ void sync() const {
    if constexpr (G > 1) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // Group sum.  A butterfly leaves each lane with the total in its own
  // association order (lane-dependent rounding), so lane 0's total is
  // broadcast: every lane must take the same accept/reject decisions and
  // build the same Householder reflectors.
  __device__ __forceinline__ double sum(double v) const {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    if constexpr (G > 1) return __shfl(v, 0, G);
    return v;
  }
  // fmax is exact, so the butterfly result is the same in every lane.
  __device__ __forceinline__ double max(double v) const {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, G));
    return v;
  }
  __device__ __forceinline__ int bcast(int v) const {
    if constexpr (G > 1) return __shfl(v, 0, G);
    return v;
  }
};

}  // namespace pdq
