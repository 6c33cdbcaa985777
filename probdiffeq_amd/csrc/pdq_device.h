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
