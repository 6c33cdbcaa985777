// Lane-per-instance kernel for the isotropic factorisation.
//
// Every lane owns one IVP instance for the whole solve: the accept/reject
// loop, checkpoint interpolation and the smoother's backward pass.  The
// per-step matrices of an instance (N x N factors, the 2N x 2N revert block,
// N x D means) are register arrays with compile-time sizes; every loop over
// them is unrolled.  The three solution slots that persist between steps
// (step_from, interp_from, proposed: scalars, mean, packed factor, backward
// conditional) live in LDS, element-major across the workgroup's 64 lanes,
// so that consecutive lanes access consecutive words; a workgroup is one
// wavefront and takes 104 KB (nu = 3) or 140 KB (nu = 4) of LDS.
// No lane ever reads data another lane wrote: the code's single-thread
// semantics are the whole story, and the same source built for the CPU
// (tests/emu) runs exactly this per-instance program.
//
// The algebra restates ssm_impl_isotropic.py:72-403 (marginalise, revert,
// merge, apply_flat, ts0 linearisation), util/cholesky_util.py:27-103
// (revert_conditional, triu_via_qr) with backend/linalg.py:8-61's R-only
// Householder QR in LAPACK's dgeqr2 sign convention, and the solver loop of
// _ivpsolve/solvers_via_adaptive_steps.py:46-375 and
// _probdiffeq/solvers.py:205-767, in the same order of operations as the
// cooperative kernel (pdq_coop.hip) and the oracle.
#pragma once

#include "pdq_args.h"
#include "pdq_device.h"

namespace pdq {

// Scheduling barrier between the phases of a step: the compiler may not move
// instructions across it, which keeps one phase's loads from being hoisted
// into the previous one (long live ranges, register spills).  No code.
#ifdef PDQ_HOST_EMU
#define PDQ_PHASE_BOUNDARY() ((void)0)
#else
#define PDQ_PHASE_BOUNDARY() __builtin_amdgcn_sched_barrier(0)
#endif


// a / b given y = 1 / b (correctly rounded): q = a y, then one FMA
// correction step against the exact residual a - b q (Markstein).  For a
// correctly rounded, normal y and a finite quotient the result is the
// correctly rounded quotient, as a division gives it, for three instructions
// once y is known; used where many quotients share one divisor.  Any other
// case (b = 0, infinities, NaN, y outside the normal range) takes the
// division itself, so special values behave exactly as IEEE division.
__device__ __forceinline__ double lane_div(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  const double res = r == 0.0 ? q : fma(r, y, q);  // exact q keeps the sign of a zero quotient
  if (__builtin_expect(!(__builtin_isnormal(y) && isfinite(res)), 0)) return a / b;
  return res;
}

// Slot scalars.
enum { LS_T = 0, LS_OS = 1, LS_NSTEPS = 2, LS_MLE_RUN = 3, LS_MLE_N = 4, LS_IDENT = 5, LS_NSC = 8 };

// Lower-triangular factors are stored packed, row by row: (i, j <= i) at
// i (i + 1) / 2 + j.
__device__ __forceinline__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// Element offsets inside one slot: scalars, mean (N x D), factor (packed),
// then the backward conditional (CondLayout order).
template <int N, int D>
struct SlotLayout {
  static constexpr int T = N * (N + 1) / 2;
  static constexpr int sc = 0;
  static constexpr int m = LS_NSC;
  static constexpr int L = m + N * D;
  static constexpr int cA = L + T;
  static constexpr int cm = cA + N * N;
  static constexpr int cL = cm + N * D;
  static constexpr int ctl = cL + T;
  static constexpr int cto = ctl + N;
  static constexpr int size = cto + N;
};

// Each lane keeps its instance's three slots in LDS, element-major across the
// workgroup's 64 lanes (element e of the lane's slot s at
// lds[(s * size + e) * 64 + lane]): consecutive lanes touch consecutive words
// and the element stride is a compile-time 512 bytes.
constexpr int kLaneBlock = 64;

// A backward conditional in latent coordinates (ssm_impl_api.py:35-132):
// A (N x N), noise mean m (N x D), noise factor L (N x N), to_latent tl (N)
// and to_observed to (N).  LCond holds one in registers; CondRef views one
// stored element-major (element e at p[e * 64]) in the blocked HBM workspace
// or in the lane's LDS staging area.  Both expose the same accessors, so the
// algebra below reads and writes either.
template <int N, int D>
struct CondLayout {
  static constexpr int T = N * (N + 1) / 2;
  static constexpr int A = 0, m = N * N, L = N * N + N * D, tl = L + T, to = tl + N, size = to + N;
};

template <int N, int D>
struct LCond {
  using CL = CondLayout<N, D>;
  double v[CL::size];
  __device__ __forceinline__ double& at(int e) { return v[e]; }
  __device__ __forceinline__ double at(int e) const { return v[e]; }
  __device__ __forceinline__ double& A(int i, int j) { return v[CL::A + i * N + j]; }
  __device__ __forceinline__ double& m(int i, int c) { return v[CL::m + i * D + c]; }
  __device__ __forceinline__ double& tl(int i) { return v[CL::tl + i]; }
  __device__ __forceinline__ double& to(int i) { return v[CL::to + i]; }
  __device__ __forceinline__ double A(int i, int j) const { return v[CL::A + i * N + j]; }
  __device__ __forceinline__ double m(int i, int c) const { return v[CL::m + i * D + c]; }
  __device__ __forceinline__ double tl(int i) const { return v[CL::tl + i]; }
  __device__ __forceinline__ double to(int i) const { return v[CL::to + i]; }
  __device__ __forceinline__ double L(int i, int j) const { return j <= i ? v[CL::L + tri(i, j)] : 0.0; }
  __device__ __forceinline__ void setL(int i, int j, double x) {
    if (j <= i) v[CL::L + tri(i, j)] = x;
  }
};

template <int N, int D>
struct CondRef {
  using CL = CondLayout<N, D>;
  double* p;
  __device__ __forceinline__ double& at(int e) const { return p[e * kLaneBlock]; }
  __device__ __forceinline__ double& A(int i, int j) const { return at(CL::A + i * N + j); }
  __device__ __forceinline__ double& m(int i, int c) const { return at(CL::m + i * D + c); }
  __device__ __forceinline__ double& tl(int i) const { return at(CL::tl + i); }
  __device__ __forceinline__ double& to(int i) const { return at(CL::to + i); }
  __device__ __forceinline__ double L(int i, int j) const { return j <= i ? at(CL::L + tri(i, j)) : 0.0; }
  __device__ __forceinline__ void setL(int i, int j, double x) const {
    if (j <= i) at(CL::L + tri(i, j)) = x;
  }
};

// View of one instance's slots in the blocked workspace.
template <int N, int D>
struct LaneWs {
  using SL = SlotLayout<N, D>;
  double* base;  // lds + lane
  __device__ __forceinline__ double& at(int slot, int e) const {
    return base[(slot * SL::size + e) * kLaneBlock];
  }
  __device__ __forceinline__ CondRef<N, D> cond(int slot) const {
    return CondRef<N, D>{base + (slot * SL::size + SL::cA) * kLaneBlock};
  }
  __device__ __forceinline__ void load_mL(int s, double* m, double* L) const {
#pragma unroll
    for (int i = 0; i < N * D; ++i) m[i] = at(s, SL::m + i);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) L[i * N + j] = j <= i ? at(s, SL::L + tri(i, j)) : 0.0;
  }
  __device__ __forceinline__ void store_mL(int s, const double* m, const double* L) const {
#pragma unroll
    for (int i = 0; i < N * D; ++i) at(s, SL::m + i) = m[i];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) at(s, SL::L + tri(i, j)) = L[i * N + j];
  }
  __device__ __forceinline__ void copy_slot(int dst, int src) const {
#pragma unroll
    for (int e = 0; e < SL::size; ++e) at(dst, e) = at(src, e);
  }
};

template <int N, int D, class Dst, class Src>
__device__ __forceinline__ void cond_copy(const Dst& d, const Src& s) {
  using CL = CondLayout<N, D>;
#pragma unroll
  for (int e = 0; e < CL::size; ++e) d.at(e) = s.at(e);
}

// ------------------------------------------------------------- algebra
// In-place R-only Householder QR of the M x C register matrix A (row-major)
// in LAPACK dgeqr2 convention (backend/linalg.py:8-37 -> geqrf): on exit the
// upper triangle of rows 0..min(M,C)-1 holds R, the strict lower part of
// those rows is zero.
//
// The reflector's norm is sqrt(alpha^2 + ss) with one FMA and one square
// root.  dgeqr2 forms it as dlapy2(alpha, dnrm2(x)) = hypot(alpha, sqrt(ss));
// both are within about an ulp of the exact norm, and the single square root
// is a much shorter dependency chain (ss itself is the same unscaled sum of
// squares in both, so the range of inputs that avoid underflow is the same).
//
// S::nz(i, r) tells, at compile time, whether row r > i can be non-zero in
// column i when reflection i is formed.  A row that cannot contributes an
// exact zero to the norm, its Householder coefficient is an exact zero and
// its update subtracts an exact zero, so skipping it leaves every result
// bit-identical to the dense dgeqr2 (for finite data) while the unrolled
// code keeps fewer values live and does fewer FLOPs.
struct DenseRows {
  static constexpr bool nz(int, int) { return true; }
};

//
// With `tau_out`, tau_out[i] receives reflection i's coefficient, 0 where the
// column needed none (for ss > 0 the coefficient is (beta - alpha) / beta with
// beta and alpha of opposite signs, never 0); the reflectors' entries stay in
// A below the diagonal of rows min(M,C).., as the updates left them.
template <int M, int C, class S = DenseRows>
__device__ __forceinline__ void lane_qr(double* A, double* tau_out = nullptr) {
  constexpr int k = M < C ? M : C;
#pragma unroll
  for (int i = 0; i < k; ++i) {
    if (M - i <= 1) break;
    if (tau_out) tau_out[i] = 0.0;
    double ss = 0.0;
#pragma unroll
    for (int r = i + 1; r < M; ++r)
      if (S::nz(i, r)) ss += A[r * C + i] * A[r * C + i];
    if (ss != 0.0) {
      const double alpha = A[i * C + i];
      // ss > 0: alpha^2 + ss, beta and alpha - beta are finite and non-zero
      // for finite data
      const double beta = -copysign(lane_sqrt_pos(fma(alpha, alpha, ss)), alpha);
      const double tau = lane_div_nz(beta - alpha, beta, lane_rcp_nz(beta));
      if (tau_out) tau_out[i] = tau;
      const double scal = lane_rcp_nz(alpha - beta);
#pragma unroll
      for (int r = i + 1; r < M; ++r)
        if (S::nz(i, r)) A[r * C + i] *= scal;
#pragma unroll
      for (int j = i + 1; j < C; ++j) {
        double w = A[i * C + j];
#pragma unroll
        for (int r = i + 1; r < M; ++r)
          if (S::nz(i, r)) w += A[r * C + i] * A[r * C + j];
        w *= tau;
        A[i * C + j] -= w;
#pragma unroll
        for (int r = i + 1; r < M; ++r)
          if (S::nz(i, r)) A[r * C + j] -= A[r * C + i] * w;
      }
      A[i * C + i] = beta;
    }
  }
#pragma unroll
  for (int i = 0; i < k; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) A[i * C + j] = 0.0;
}

// [dense N x N; upper-triangular N x N] stacks (sum_of_sqrtm_factors of a
// product factor and a Cholesky factor): bottom row N + q is zero in the
// columns before q, and reflections never fill it in.
template <int N>
struct TriStackRows {
  static constexpr bool nz(int i, int r) { return r < N || r - N <= i; }
};

// The revert block [[R_YX, 0], [R_X_F, R_X]] with R_YX upper triangular:
// below the pivot, the left block's top rows stay zero; after the left
// block, the bottom-right block is dense.
template <int N>
struct RevertRows {
  static constexpr bool nz(int i, int r) { return i >= N || r >= N; }
};

// The update block [[damp, 0], [(H L)^T, L^T]] of ts0 (H selects derivative
// 1): column 0 has non-zeros in rows 0..2 only, and after two reflections
// the block is upper triangular.
struct UpdateRows {
  static constexpr bool nz(int i, int r) { return i <= 1 && r <= 2; }
};

// fx.marginalise's stacked column [(H L)^T; damp] for a prior factor L =
// |p| (s Q) with Q lower triangular: rows 0, 1 and the damping row.
template <int N>
struct FxColumnRows {
  static constexpr bool nz(int, int r) { return r <= 1 || r == N; }
};

template <int N, int i = 0>
__device__ __forceinline__ void lane_precon(double dt, double* P, double* PI) {
  if constexpr (i < N) {
    constexpr int pw = N - 1 - i;
    const double s = factorial_d(pw);
    P[i] = lane_div_nz(lane_ipow<pw>(dt), s, 1.0 / s);  // 1 / s correctly rounded at compile time
    PI[i] = lane_ipow<-pw>(dt) * s;
    lane_precon<N, i + 1>(dt, P, PI);
  }
}

// Isotropic noise scale (ssm_impl_isotropic.py:375-376): (sqrt|dt| * base) * s.
__device__ __forceinline__ double lane_noise_scale(const pdq_config& cfg, double dt, double s) {
  return sqrt(fabs(dt)) * cfg.base_output_scale * s;
}

// The IWP transition in preconditioned coordinates (utilities.py:57-71):
// A[i][j] = C(nu - i, nu - j) for j >= i, else 0.  These are small exact
// integers (the values pdq_iwp.h computes, by the same recurrence), known here
// at compile time: in the unrolled products a zero entry's term is dropped
// (it adds an exact zero) and a unit entry's product is the other factor, so
// for finite data the results are bit-identical to the generic product.
constexpr double lane_binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
  return r;
}

template <int N>
constexpr double lane_iwp_a(int i, int j) {
  return j < i ? 0.0 : lane_binom(N - 1 - i, N - 1 - j);
}

// The process-noise factor Q is lower triangular; its upper entries are
// (signed) zeros that every consumer treats as structural zeros.
template <int N>
__device__ __forceinline__ double lane_iwp_q(const double* __restrict__ Q, int i, int j) {
  return j <= i ? Q[i * N + j] : 0.0;
}

// apply_flat of the transition to a mean (ssm_impl_isotropic.py:75-79).
template <int N, int D>
__device__ __forceinline__ void lane_apply_flat_mean(const double* P, const double* PI, const double* x, double* out) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(i, k);
        if (a != 0.0) acc += (P[i] * a) * (PI[k] * x[k * D + col]);
      }
      out[i * D + col] = acc + 0.0;
    }
}

// marginalise of the transition (ssm_impl_isotropic.py:81-89).
template <int N, int D>
__device__ __forceinline__ void lane_marginalise_prior(const double* __restrict__ Q,
                                                       const double* P, const double* PI, double s, const double* Xm,
                                                       const double* XL, double* Ym, double* YL) {
  double big[2 * N * N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(col, k);
        if (a != 0.0) acc += a * (PI[k] * XL[k * N + r]);
      }
      big[r * N + col] = acc;
      big[(N + r) * N + col] = s * lane_iwp_q<N>(Q, col, r);
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(i, k);
        if (a != 0.0) acc += a * (PI[k] * Xm[k * D + col]);
      }
      Ym[i * D + col] = P[i] * (acc + 0.0);
    }
  lane_qr<2 * N, N, TriStackRows<N>>(big);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) YL[i * N + j] = fabs(P[i]) * big[j * N + i];
}

// revert of the transition (ssm_impl_isotropic.py:107-133 with
// cholesky_util.py:27-82): observed (Ym, YL) and backward conditional bw
// (written to any conditional storage).
template <int N, int D, class Cond>
// With `ident_rescale` the conditional is stored already merged behind an
// identity (lane_merge with an identity `self`: rows scaled by to_observed,
// to_observed = 1), bitwise as the revert followed by that merge.
__device__ __forceinline__ void lane_revert_prior(const double* __restrict__ Q,
                                                  const double* P, const double* PI, double s, const double* Xm,
                                                  const double* XL, double* Ym, double* YL, Cond&& bw,
                                                  bool ident_rescale = false) {
  constexpr int n2 = 2 * N;
  double big[n2 * n2];
  double tmp[N * D];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      big[r * n2 + col] = s * lane_iwp_q<N>(Q, col, r);
      big[r * n2 + N + col] = 0.0;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(col, k);
        if (a != 0.0) acc += a * (fabs(PI[k]) * XL[k * N + r]);
      }
      big[(N + r) * n2 + col] = acc;
      big[(N + r) * n2 + N + col] = fabs(PI[col]) * XL[col * N + r];
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(i, k);
        if (a != 0.0) acc += a * (PI[k] * Xm[k * D + col]);
      }
      tmp[i * D + col] = acc + 0.0;
    }
  lane_qr<n2, n2, RevertRows<N>>(big);
  // gain = solve_triu(R_Y, R12)^T
  double G[N * N], rdiag[N];
#pragma unroll
  for (int i = 0; i < N; ++i) rdiag[i] = lane_rcp_nz(big[i * n2 + i]);
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      double sacc = big[i * n2 + N + j];
#pragma unroll
      for (int c = i + 1; c < N; ++c) sacc -= big[i * n2 + c] * G[j * N + c];
      G[j * N + i] = lane_div(sacc, big[i * n2 + i], rdiag[i]);
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += G[i * N + k] * tmp[k * D + col];
      const double mv = PI[i] * Xm[i * D + col] - acc;
      bw.m(i, col) = ident_rescale ? (1.0 * lane_rcp_nz(PI[i])) * mv + 0.0 : mv;
      Ym[i * D + col] = P[i] * tmp[i * D + col];
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double to = lane_rcp_nz(PI[i]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      bw.A(i, j) = ident_rescale ? (1.0 * to) * G[i * N + j] : G[i * N + j];
      const double lv = big[(N + j) * n2 + N + i];
      bw.setL(i, j, ident_rescale ? fabs(1.0 * to) * lv : lv);
      YL[i * N + j] = fabs(P[i]) * big[j * n2 + i];
    }
    bw.tl(i) = lane_rcp_nz(P[i]);
    bw.to(i) = ident_rescale ? 1.0 : to;
  }
}

// lane_revert_prior in two halves, for the fixed-point smoother's adaptive
// loop: the left half gives the prediction (what the step needs) and goes on
// to the revert block's right columns, the right half finishes the backward
// conditional (what only the merges and checkpoints read), so that a second
// wavefront can do it beside the step's update.  Reflection i < N of the
// revert block has its non-zeros in row i and rows N..2N-1 (RevertRows), and
// a reflection updates each column from that column and the reflector alone.
// So the left half's QR of the left N columns gives lane_revert_prior's R_Y,
// and its reflections of the right N columns (in lane_qr's order) and the QR
// of their rows N..2N-1 give R12 and R22: bitwise those of
// lane_revert_prior, as are the right half's gain and conditional.
//
// Between the halves the stage below sits in storage the caller provides
// (StageRef: the conditional's own storage, which the right half then
// overwrites, and a tail beyond it): R12, the upper triangles of R22 and
// R_Y, and the preconditioner P, PI.
template <int N, int D>
struct RevertStage {
  static constexpr int T = N * (N + 1) / 2;
  static constexpr int r12 = 0, r22 = N * N, r = r22 + T, P = r + T, PI = P + N, size = PI + N;
  static constexpr int tail = size > CondLayout<N, D>::size ? size - CondLayout<N, D>::size : 0;
};

// The stage's storage: element e < the conditional's size in the
// conditional (stride kLaneBlock), the rest in `x` (stride `xs`).
template <int N, int D>
struct StageRef {
  double* c;
  double* x;
  int xs;
  __device__ __forceinline__ double& at(int e) const {
    constexpr int cs = CondLayout<N, D>::size;
    return e < cs ? c[e * kLaneBlock] : x[(e - cs) * xs];
  }
};

template <int N, int D>
__device__ __forceinline__ void lane_revert_left(const double* __restrict__ Q, const double* P, const double* PI,
                                                 double s, const double* Xm, const double* XL, double* Ym, double* YL,
                                                 const StageRef<N, D>& st) {
  using RS = RevertStage<N, D>;
  double big[2 * N * N], tau[N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      big[r * N + col] = s * lane_iwp_q<N>(Q, col, r);
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(col, k);
        if (a != 0.0) acc += a * (fabs(PI[k]) * XL[k * N + r]);
      }
      big[(N + r) * N + col] = acc;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(i, k);
        if (a != 0.0) acc += a * (PI[k] * Xm[k * D + col]);
      }
      Ym[i * D + col] = P[i] * (acc + 0.0);
    }
  lane_qr<2 * N, N, RevertRows<N>>(big, tau);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      YL[i * N + j] = fabs(P[i]) * big[j * N + i];
      if (j >= i) st.at(RS::r + tri(j, i)) = big[i * N + j];
    }
  // the right columns: zeros on top, (|PI| XL)^T below, then lane_qr's
  // column update for reflections 0..N-1
  double rt[2 * N * N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      rt[r * N + col] = 0.0;
      rt[(N + r) * N + col] = fabs(PI[col]) * XL[col * N + r];
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (tau[i] != 0.0) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double w = rt[i * N + j];
#pragma unroll
        for (int r = 0; r < N; ++r) w += big[(N + r) * N + i] * rt[(N + r) * N + j];
        w *= tau[i];
        rt[i * N + j] -= w;
#pragma unroll
        for (int r = 0; r < N; ++r) rt[(N + r) * N + j] -= big[(N + r) * N + i] * w;
      }
    }
  }
  lane_qr<N, N>(rt + N * N);  // reflections N..2N-1: the dense bottom-right block
#pragma unroll
  for (int e = 0; e < N * N; ++e) st.at(RS::r12 + e) = rt[e];
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i; j < N; ++j) st.at(RS::r22 + tri(j, i)) = rt[(N + i) * N + j];
    st.at(RS::P + i) = P[i];
    st.at(RS::PI + i) = PI[i];
  }
}

// The right half: Xm is the reverted state's mean, as the left half had it;
// the stage is read before `bw`, the conditional, is written (with
// `ident_rescale` as in lane_revert_prior).
template <int N, int D, class Cond>
__device__ __forceinline__ void lane_revert_right(const double* Xm, const StageRef<N, D>& st, Cond&& bw,
                                                  bool ident_rescale) {
  using RS = RevertStage<N, D>;
  double rt[2 * N * N], R[N * N], P[N], PI[N];
#pragma unroll
  for (int e = 0; e < N * N; ++e) rt[e] = st.at(RS::r12 + e);
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      R[i * N + j] = j >= i ? st.at(RS::r + tri(j, i)) : 0.0;
      rt[(N + i) * N + j] = j >= i ? st.at(RS::r22 + tri(j, i)) : 0.0;
    }
    P[i] = st.at(RS::P + i);
    PI[i] = st.at(RS::PI + i);
  }
  double tmp[N * D];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(i, k);
        if (a != 0.0) acc += a * (PI[k] * Xm[k * D + col]);
      }
      tmp[i * D + col] = acc + 0.0;
    }
  double G[N * N], rdiag[N];
#pragma unroll
  for (int i = 0; i < N; ++i) rdiag[i] = lane_rcp_nz(R[i * N + i]);
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      double sacc = rt[i * N + j];
#pragma unroll
      for (int c = i + 1; c < N; ++c) sacc -= R[i * N + c] * G[j * N + c];
      G[j * N + i] = lane_div(sacc, R[i * N + i], rdiag[i]);
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += G[i * N + k] * tmp[k * D + col];
      const double mv = PI[i] * Xm[i * D + col] - acc;
      bw.m(i, col) = ident_rescale ? (1.0 * lane_rcp_nz(PI[i])) * mv + 0.0 : mv;
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double to = lane_rcp_nz(PI[i]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      bw.A(i, j) = ident_rescale ? (1.0 * to) * G[i * N + j] : G[i * N + j];
      const double lv = rt[(N + j) * N + i];
      bw.setL(i, j, ident_rescale ? fabs(1.0 * to) * lv : lv);
    }
    bw.tl(i) = lane_rcp_nz(P[i]);
    bw.to(i) = ident_rescale ? 1.0 : to;
  }
}

// merge (ssm_impl_isotropic.py:91-105): r = self o other.  r may be the same
// storage as `other` (every part of r is written after the parts of other it
// depends on have been read).  With an identity `self` the general formula
// reduces exactly to a rescaling (the QR of an upper-triangular factor
// stacked on zeros is that factor).
template <int N, int D, class CS, class CO, class CR>
__device__ __forceinline__ void lane_merge(bool self_ident, const CS& sf, const CO& ot, const CR& r) {
  if (self_ident) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double T = 1.0 * ot.to(i);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        r.A(i, j) = T * ot.A(i, j);
        r.setL(i, j, fabs(T) * ot.L(i, j));
      }
#pragma unroll
      for (int col = 0; col < D; ++col) r.m(i, col) = (1.0 * ot.to(i)) * ot.m(i, col) + 0.0;
      r.tl(i) = ot.tl(i);
      r.to(i) = 1.0;
    }
    return;
  }
  double T[N], sA[N * N];
#pragma unroll
  for (int k = 0; k < N; ++k) T[k] = sf.tl(k) * ot.to(k);
#pragma unroll
  for (int e = 0; e < N * N; ++e) sA[e] = sf.A(e / N, e % N);
  {
    double oA[N * N], rA[N * N];
#pragma unroll
    for (int e = 0; e < N * N; ++e) oA[e] = ot.A(e / N, e % N);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) acc += sA[i * N + k] * (T[k] * oA[k * N + j]);
        rA[i * N + j] = acc;
      }
#pragma unroll
    for (int e = 0; e < N * N; ++e) r.A(e / N, e % N) = rA[e];
  }
  {
    double om[N * D], rm[N * D];
#pragma unroll
    for (int e = 0; e < N * D; ++e) om[e] = ot.m(e / D, e % D);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < D; ++col) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) acc += sA[i * N + k] * (T[k] * om[k * D + col]);
        rm[i * D + col] = acc + sf.m(i, col);
      }
#pragma unroll
    for (int e = 0; e < N * D; ++e) r.m(e / D, e % D) = rm[e];
  }
  {
    double big[2 * N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double acc2 = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) acc2 += sA[i * N + k] * (fabs(T[k]) * ot.L(k, j));
        big[j * N + i] = acc2;
        big[(N + j) * N + i] = sf.L(i, j);
      }
    lane_qr<2 * N, N, TriStackRows<N>>(big);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) r.setL(i, j, big[j * N + i]);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    r.tl(i) = ot.tl(i);
    r.to(i) = sf.to(i);
  }
}

// marginalise of a backward conditional (estimators_and_losses.py:156-178).
template <int N, int D, class Cond>
__device__ __forceinline__ void lane_marginalise_cond(const Cond& c, const double* Xm, const double* XL, double* Ym,
                                                      double* YL) {
  double big[2 * N * N], tl[N];
#pragma unroll
  for (int k = 0; k < N; ++k) tl[k] = c.tl(k);
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += c.A(col, k) * (tl[k] * XL[k * N + r]);
      big[r * N + col] = acc;
      big[(N + r) * N + col] = c.L(col, r);
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += c.A(i, k) * (tl[k] * Xm[k * D + col]);
      Ym[i * D + col] = c.to(i) * (acc + c.m(i, col));
    }
  lane_qr<2 * N, N, TriStackRows<N>>(big);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) YL[i * N + j] = fabs(c.to(i)) * big[j * N + i];
}

// The linearisation at the mean Xm: the observed residual is
// h0 x_0 + x_1 + b (the 1 x N linear operator [h0, 1, 0...]).
// * ts0 (ssm_impl_isotropic.py:298-317): b = -f(u(t)) and the operator
//   selects derivative 1 (h0 is not applied).
// * ts1 (IsotropicResidual.linearize, ssm_impl_isotropic.py:321-355) with
//   jacobian_materialize's exact trace (jacobians.py:100-106): the trace over
//   d of d r / d x for r = x_1 - f(x_0) is [-tr J_f, d, 0...], divided by d,
//   and b = r(Xm) - [h0, 1] Xm.  With jacobian_monte_carlo_rev or _fwd (:230-248, :151-169)
//   tr J_f is the probe mean of v^T J_f v (probe set `set`); the entry of x_1,
//   the mean of v^T v = d, stays exact.
template <int N, int D>
__device__ __forceinline__ void lane_linearize(const pdq_config& cfg, const double* __restrict__ prm, const double* Xm,
                                               double t, double* b, double& h0, uint32_t set = 0) {
  double y[D];
#pragma unroll
  for (int j = 0; j < D; ++j) y[j] = Xm[j];
  if (cfg.constraint != PDQ_TS1) {
    h0 = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) b[j] = -vf_component(cfg.vector_field, D, y, prm, t, j);
    return;
  }
  double tr = 0.0;
  if (cfg.jacobian != PDQ_JAC_EXACT) {
    const bool fwd = cfg.jacobian == PDQ_JAC_MONTE_CARLO_FWD;
    for (int j = 0; j < D; ++j)
      tr += -vf_mc_diag(cfg.vector_field, D, y, prm, j, cfg.jac_seed, set, cfg.jac_probes, fwd);
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) tr += -vf_jac(cfg.vector_field, D, y, prm, j, j);
  }
  h0 = tr / (double)D;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const double r = Xm[1 * D + j] - vf_component(cfg.vector_field, D, y, prm, t, j);
    b[j] = r - (h0 * Xm[j] + Xm[1 * D + j]);
  }
}

// The linear operator applied to row-major N x C rows: h0 X[0, c] + X[1, c]
// (ts1), or X[1, c] (ts0).
template <int C>
__device__ __forceinline__ double lane_obs_row(const pdq_config& cfg, double h0, const double* X, int c) {
  return cfg.constraint == PDQ_TS1 ? h0 * X[c] + X[1 * C + c] : X[1 * C + c];
}

// fx.marginalise(u): observed mean om (D) and factor oL (1 x 1).  The
// column (H L)^T has the sparsity of row 1 of L for both linearisations.
template <int N, int D>
__device__ __forceinline__ void lane_marginalise_fx(const pdq_config& cfg, const double* b, double h0,
                                                    const double* Xm, const double* XL, double* om, double& oL) {
  double col[N + 1];
#pragma unroll
  for (int r = 0; r < N; ++r) col[r] = lane_obs_row<N>(cfg, h0, XL, r);
  col[N] = cfg.damp;
#pragma unroll
  for (int c = 0; c < D; ++c) om[c] = lane_obs_row<D>(cfg, h0, Xm, c) + b[c];
  lane_qr<N + 1, 1, FxColumnRows<N>>(col);
  oL = col[0];
}

// residual_whitened_rms against zero data (ssm_impl_isotropic.py:203-207).
template <int D>
__device__ __forceinline__ double lane_rms_zero(const double* om, double oL) {
  double ss = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const double w = om[c] / oL;
    ss += w * w;
  }
  return sqrt(ss) / sqrt((double)D);
}

// fx.revert(pred) + apply_flat(0) (ssm_impl_api.py:112-132): the update.
template <int N, int D>
__device__ __forceinline__ void lane_update(const pdq_config& cfg, const double* b, double h0, const double* Pm,
                                            const double* PL, double* Xm, double* XL, double* mo, double& oL_out) {
  constexpr int n2 = N + 1;
  double big[n2 * n2];
#pragma unroll
  for (int r = 0; r < n2; ++r)
#pragma unroll
    for (int col = 0; col < n2; ++col) {
      double v;
      if (r < 1) v = col < 1 ? cfg.damp : 0.0;
      else v = col < 1 ? lane_obs_row<N>(cfg, h0, PL, r - 1) : PL[(col - 1) * N + (r - 1)];
      big[r * n2 + col] = v;
    }
#pragma unroll
  for (int c = 0; c < D; ++c) mo[c] = lane_obs_row<D>(cfg, h0, Pm, c) + b[c];
  lane_qr<n2, n2, UpdateRows>(big);
  double gain[N];
  const double r0 = lane_rcp_nz(big[0]);
#pragma unroll
  for (int j = 0; j < N; ++j) gain[j] = lane_div(big[1 + j], big[0], r0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double acc = 0.0 + gain[i] * mo[c];
      Xm[i * D + c] = Pm[i * D + c] - acc;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) XL[i * N + j] = big[(1 + j) * n2 + 1 + i];
  oL_out = big[0];
}

// The error power of an isotropic error estimate (estimate_error_norm of
// error_residual_std, solvers.py:925-996, and error_state_std, :1041-1098):
// err is the calibrated std (one scalar for the isotropic factorisation),
// `row` the derivative whose means give the reference, and nord the order n
// in error_abs = err dt^n / n!.
template <int N, int D>
__device__ __forceinline__ double lane_error_power(const pdq_config& cfg, double err, int row, const double* prev_m,
                                                   const double* prop_m, double dt, int nord) {
  const double e = nord == 1 ? err * dt / 1.0 : err * pow(dt, (double)nord) / factorial_d(nord);
  double ref[D];
#pragma unroll
  for (int k = 0; k < D; ++k) ref[k] = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r)  // the reference row, selected without indexing registers by a run-time row
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (r == row) ref[k] = fmax(fabs(prev_m[r * D + k]), fabs(prop_m[r * D + k]));
  double norm;
  if (cfg.error_norm == PDQ_NORM_SCALE_THEN_RMS) {
    double ss = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double rr = e / (cfg.atol + cfg.rtol * fabs(ref[k]));
      ss += rr * rr;
    }
    norm = sqrt(ss) / sqrt((double)D);
  } else {
    const double sa = 0.0 + e * e;
    double sr = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) sr += ref[k] * ref[k];
    norm = (sqrt(sa) / sqrt(1.0)) / (cfg.atol + cfg.rtol * (sqrt(sr) / sqrt((double)D)));
  }
  return lane_root_inv<N>(norm);
}

// error_residual_std's calibrated std of the observed residual: the
// isotropic std (vector norm) of the rescaled 1 x 1 factor sigma oL.
__device__ __forceinline__ double lane_residual_err(double sigma, double oL) {
  const double alpha = sigma * oL;
  return sqrt(alpha * alpha + 0.0);
}

// error_state_std's calibrated std (solvers.py:1070-1084): condition the
// output-scale-1 extrapolation (Um, UL) on the linearisation b
// (bayes_rule_and_residual_whitened_rms_tree), then scale the std of
// derivative idx of the result by the residual's output scale.
template <int N, int D>
__device__ __forceinline__ double lane_state_err(const pdq_config& cfg, const double* b, double h0, const double* Um,
                                                 const double* UL, int idx) {
  double Xm[N * D], XL[N * N], mo[D], oL;
  lane_update<N, D>(cfg, b, h0, Um, UL, Xm, XL, mo, oL);
  const double os = lane_rms_zero<D>(mo, oL);
  double std_idx = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double ss = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) ss += XL[r * N + j] * XL[r * N + j];
    if (r == idx) std_idx = sqrt(ss);
  }
  return os * std_idx;
}

// controllers.py:66-84 / :24-63
__device__ __forceinline__ double lane_control(const pdq_config& cfg, double dt, double& cstate, double ep) {
  if (cfg.control == PDQ_CONTROL_INTEGRAL) {
    const double ratio = cfg.safety * ep;
    return fmax(cfg.factor_min, fmin(ratio, cfg.factor_max)) * dt;
  }
  const double gi = pow(ep, cfg.exponent_integral);
  const double gp = pow(ep / cstate, cfg.exponent_proportional);
  const double ratio = cfg.safety * gi * gp;
  const double sf = fmax(cfg.factor_min, fmin(ratio, cfg.factor_max));
  if (ep >= 1.0) cstate = ep;
  return sf * dt;
}

// ------------------------------------------------------------ instance
template <int N, int D>
struct LaneCtx {
  const LArgs* a;
  const double* prm;
  LaneWs<N, D> w;
  int inst;
  uint32_t lin_set;  // the step's probe set of the Monte Carlo Jacobian handlers (pdq_lin_set)
  double* stage_tail;  // the revert stage's tail (StageRef), element stride stage_stride
  int stage_stride;
};

// The three parts of a step attempt: solver_dynamic.step (solvers.py:549-599),
// solver.step (:702-733) and solver_mle.step (:395-437) with the residual
// error estimate, from slot `from` to slot `to`.
//
// (1) Calibration (solver_dynamic only): extrapolate the mean with output
// scale 1, linearise there, sigma = the whitened residual's RMS.  `err` is
// the error estimate's calibrated std, taken here where its operands are
// live: the error's own re-linearisation (re_linearize_before_error) would
// linearise at this same extrapolation, so it reproduces b exactly.
template <int N, int D>
__device__ __forceinline__ void lane_calibrate(const LaneCtx<N, D>& c, const double* Sm, const double* P,
                                               const double* PI, double h, double tn, double* b, double& h0,
                                               double& sigma, double& err, bool estimate) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const double s1 = lane_noise_scale(cfg, h, 1.0);
  double Um[N * D], UL[N * N];
  lane_apply_flat_mean<N, D>(P, PI, Sm, Um);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) UL[i * N + j] = fabs(P[i]) * (s1 * lane_iwp_q<N>(Q, i, j));
  lane_linearize<N, D>(cfg, c.prm, Um, tn, b, h0, c.lin_set);
  double om[D], oL;
  lane_marginalise_fx<N, D>(cfg, b, h0, Um, UL, om, oL);
  sigma = lane_rms_zero<D>(om, oL);
  err = 0.0;
  if (estimate)
    err = cfg.error_estimator == PDQ_ERROR_STATE_STD
              ? lane_state_err<N, D>(cfg, b, h0, Um, UL, cfg.error_derivative_idx)
              : lane_residual_err(sigma, oL);
}

// (2) The prediction, shared by step attempts and checkpoint interpolation:
// the filter marginalises the transition; the smoothers revert it into the
// conditional of slot `bws`, and the fixed-point smoother then merges the
// conditional of slot `mself` (identity when `mident`) in front of it,
// writing slot `mres` (which may be `bws` or `mself`).
template <int N, int D>
__device__ __forceinline__ void lane_predict(const LaneCtx<N, D>& c, const double* P, const double* PI, double s,
                                             const double* Xm, const double* XL, double* Pm, double* PL, int bws,
                                             bool merge, bool mident, int mself, int mres) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  if (cfg.strategy == PDQ_FILTER) {
    lane_marginalise_prior<N, D>(Q, P, PI, s, Xm, XL, Pm, PL);
    return;
  }
  lane_revert_prior<N, D>(Q, P, PI, s, Xm, XL, Pm, PL, c.w.cond(bws));
  PDQ_PHASE_BOUNDARY();
  if (merge) lane_merge<N, D>(mident, c.w.cond(mself), c.w.cond(bws), c.w.cond(mres));
  PDQ_PHASE_BOUNDARY();
}

// (3) The update, the MLE running scale, the error estimate and the proposed
// solution's slot.  b, h0, sigma and err come from (1) for solver_dynamic.
// Returns error_power (meaningful with `estimate`).
template <int N, int D>
__device__ __forceinline__ double lane_finish_step(const LaneCtx<N, D>& c, int from, int to, double h, double tn,
                                                   const double* P, const double* PI, const double* Sm,
                                                   const double* Pm, const double* PL, double* b, double h0,
                                                   double sigma, double err, bool estimate) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const LaneWs<N, D>& w = c.w;
  double Rm[N * D], RL[N * N], mo[D], oLu, ep = 0.0;
  const bool state_std = cfg.error_estimator == PDQ_ERROR_STATE_STD;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    // solvers.py:576-582: linearise again at the calibrated prediction; the
    // step reports its first linearisation (:597), which the error uses
    double bu[D], hu = h0;
#pragma unroll
    for (int k = 0; k < D; ++k) bu[k] = b[k];
    if (cfg.re_linearize_after_calibration) lane_linearize<N, D>(cfg, c.prm, Pm, tn, bu, hu, c.lin_set + 1);
    lane_update<N, D>(cfg, bu, hu, Pm, PL, Rm, RL, mo, oLu);
  } else {
    const double s1 = lane_noise_scale(cfg, h, 1.0);
    lane_linearize<N, D>(cfg, c.prm, Pm, tn, b, h0, c.lin_set);
    if (estimate) {
      // the error's calibrated std, before the update so that its operands
      // and the posterior are not live together
      double Um[N * D], UL[N * N], be[D], he = h0;
      lane_apply_flat_mean<N, D>(P, PI, Sm, Um);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) UL[i * N + j] = fabs(P[i]) * (s1 * lane_iwp_q<N>(Q, i, j));
      // the step's linearisation, or a new one at the extrapolation
      // (re_linearize_before_error, solvers.py:951-956)
#pragma unroll
      for (int k = 0; k < D; ++k) be[k] = b[k];
      if (cfg.error_re_linearize) lane_linearize<N, D>(cfg, c.prm, Um, tn, be, he, c.lin_set);
      if (state_std) {
        err = lane_state_err<N, D>(cfg, be, he, Um, UL, cfg.error_derivative_idx);
      } else {
        double om[D], oLe;
        lane_marginalise_fx<N, D>(cfg, be, he, Um, UL, om, oLe);
        err = lane_residual_err(lane_rms_zero<D>(om, oLe), oLe);
      }
    }
    const bool mle = cfg.solver == PDQ_SOLVER_MLE;
    lane_update<N, D>(cfg, b, h0, Pm, PL, Rm, RL, mo, oLu);
    if (mle) {
      // bayes_rule_and_residual_whitened_rms_tree: observed = (mo, R_Y^T)
      const double nt = lane_rms_zero<D>(mo, oLu);
      const double nd = w.at(from, LS_MLE_N);
      const double x1 = sqrt(nd / (nd + 1)) * w.at(from, LS_MLE_RUN);
      const double x2 = sqrt(1 / (nd + 1)) * nt;
      w.at(to, LS_MLE_RUN) = hypot(x1, x2);
      w.at(to, LS_MLE_N) = nd + 1;
    }
    sigma = mle ? w.at(from, LS_OS) : 1.0;
  }
  if (estimate) {
    const int err_row = state_std ? cfg.error_derivative_idx : 0;
    const int nord = (state_std ? cfg.error_derivative_idx : 1) + (cfg.error_per_unit_step ? 1 : 0);
    ep = lane_error_power<N, D>(cfg, err, err_row, Sm, Rm, h, nord);
  }
  w.store_mL(to, Rm, RL);
  w.at(to, LS_T) = tn;
  w.at(to, LS_OS) = sigma;
  w.at(to, LS_NSTEPS) = w.at(from, LS_NSTEPS) + 1.0;
  w.at(to, LS_IDENT) = 0.0;
  if (cfg.solver != PDQ_SOLVER_MLE) {
    w.at(to, LS_MLE_RUN) = 0.0;
    w.at(to, LS_MLE_N) = 0.0;
  }
  return ep;
}

// One step attempt from slot `from` to slot `to` (the three parts in order).
template <int N, int D>
__device__ __forceinline__ double lane_step(const LaneCtx<N, D>& c, int from, int to, double h, bool estimate) {
  const pdq_config& cfg = c.a->cfg;
  const LaneWs<N, D>& w = c.w;
  double Sm[N * D], SLf[N * N], P[N], PI[N], b[D], Pm[N * D], PL[N * N];
  w.load_mL(from, Sm, SLf);
  const double tn = w.at(from, LS_T) + h;
  lane_precon<N>(h, P, PI);
  double sigma = 1.0, err = 0.0, h0 = 0.0;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) lane_calibrate<N, D>(c, Sm, P, PI, h, tn, b, h0, sigma, err, estimate);
  PDQ_PHASE_BOUNDARY();
  lane_predict<N, D>(c, P, PI, lane_noise_scale(cfg, h, sigma), Sm, SLf, Pm, PL, to,
                     cfg.strategy == PDQ_FIXEDPOINT, w.at(from, LS_IDENT) != 0.0, from, to);
  return lane_finish_step<N, D>(c, from, to, h, tn, P, PI, Sm, Pm, PL, b, h0, sigma, err, estimate);
}

// ---------------------------------------------------------------- outputs
template <int N, int D>
__device__ __forceinline__ void lane_write_normal(const double* m, const double* L, double* gm, double* gL) {
  if (gm)
#pragma unroll
    for (int i = 0; i < N * D; ++i) gm[i] = m[i];
  if (gL)
#pragma unroll
    for (int i = 0; i < N * N; ++i) gL[i] = L[i];
}

// Backward conditional j of the instance's output (identity when `ident`).
template <int N, int D, class Cond>
__device__ __forceinline__ void lane_write_cond(const LaneCtx<N, D>& c, bool ident, const Cond& s, int j) {
  const pdq_solution& o = c.a->out;
  const size_t idx = (size_t)c.inst * (c.a->K - 1) + j;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int k = 0; k < N; ++k) {
      o.cond_A[idx * N * N + i * N + k] = ident ? (i == k ? 1.0 : 0.0) : s.A(i, k);
      o.cond_chol[idx * N * N + i * N + k] = ident ? 0.0 : s.L(i, k);
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int col = 0; col < D; ++col) o.cond_mean[idx * N * D + i * D + col] = ident ? 0.0 : s.m(i, col);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    o.cond_tl[idx * N + i] = ident ? 1.0 : s.tl(i);
    o.cond_to[idx * N + i] = ident ? 1.0 : s.to(i);
  }
}

// Writes checkpoint k: the filter's marginal, or the smoothers' filtering
// marginal (and, with `cond`, the backward conditional that leads to it).
template <int N, int D, class Cond>
__device__ __forceinline__ void lane_write_checkpoint(const LaneCtx<N, D>& c, int k, double t, const double* m,
                                                      const double* L, double os, double nsteps, const Cond* cond,
                                                      bool cond_ident) {
  const pdq_solution& o = c.a->out;
  const int K = c.a->K;
  const size_t row = (size_t)c.inst * K + k;
  const pdq_config& cfg = c.a->cfg;
  if (o.t) o.t[row] = t;
  if (o.output_scale && cfg.solver == PDQ_SOLVER_DYNAMIC) o.output_scale[row] = os;
  if (o.num_steps && k > 0) o.num_steps[(size_t)c.inst * (K - 1) + k - 1] = (int64_t)nsteps;
  if (cfg.strategy != PDQ_FILTER) {
    lane_write_normal<N, D>(m, L, o.filt_mean ? o.filt_mean + row * N * D : nullptr,
                            o.filt_chol ? o.filt_chol + row * N * N : nullptr);
    if (k > 0 && cond) lane_write_cond<N, D>(c, cond_ident, *cond, k - 1);
  } else {
    lane_write_normal<N, D>(m, L, o.mean ? o.mean + row * N * D : nullptr, o.chol ? o.chol + row * N * N : nullptr);
  }
}

// Checkpoint k from slot s (its own marginal and backward conditional).
template <int N, int D>
__device__ __forceinline__ void lane_write_slot(const LaneCtx<N, D>& c, int k, int s) {
  double m[N * D], L[N * N];
  c.w.load_mL(s, m, L);
  const CondRef<N, D> cond = c.w.cond(s);
  lane_write_checkpoint<N, D>(c, k, c.w.at(s, LS_T), m, L, c.w.at(s, LS_OS), c.w.at(s, LS_NSTEPS), &cond,
                              c.w.at(s, LS_IDENT) != 0.0);
}

// ProbabilisticSolver.interpolate_fwd (solvers.py:205-269): from = interp_from,
// to = step_from; writes checkpoint k and updates both slots.
template <int N, int D>
__device__ __forceinline__ void lane_interp_beyond(const LaneCtx<N, D>& c, int k, double t1, int from, int to) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const LaneWs<N, D>& w = c.w;
  const double os = w.at(to, LS_OS);
  const double dt0 = t1 - w.at(from, LS_T);
  const double dt1 = w.at(to, LS_T) - t1;
  const double nsteps = w.at(to, LS_NSTEPS);
  double Fm[N * D], FL[N * N], P[N], PI[N], Pm[N * D], PL[N * N];
  w.load_mL(from, Fm, FL);
  lane_precon<N>(dt0, P, PI);
  const double s0 = lane_noise_scale(cfg, dt0, os);
  if (cfg.strategy == PDQ_FILTER) {
    lane_marginalise_prior<N, D>(Q, P, PI, s0, Fm, FL, Pm, PL);
    lane_write_checkpoint<N, D, CondRef<N, D>>(c, k, t1, Pm, PL, os, nsteps, nullptr, false);
    w.store_mL(from, Pm, PL);
    w.at(from, LS_T) = t1;
    return;
  }
  // extrapolate t0 -> t: ext_t = (marginal Pm/PL, conditional ext).  ext is
  // built in interp_from's conditional, which it replaces: fixed point
  // ext = interp_from.cond o bw (bw staged in step_from's conditional, which
  // is rewritten below); fixed interval ext = bw.
  if (cfg.strategy == PDQ_FIXEDPOINT) {
    lane_revert_prior<N, D>(Q, P, PI, s0, Fm, FL, Pm, PL, w.cond(to));
    PDQ_PHASE_BOUNDARY();
    lane_merge<N, D>(w.at(from, LS_IDENT) != 0.0, w.cond(from), w.cond(to), w.cond(from));
  } else {
    lane_revert_prior<N, D>(Q, P, PI, s0, Fm, FL, Pm, PL, w.cond(from));
  }
  PDQ_PHASE_BOUNDARY();
  // extrapolate t -> t1 from ext_t's marginal; only the backward conditional
  // is kept, as step_from's new conditional (fixed point: identity o bw)
  lane_precon<N>(dt1, P, PI);
  const double s1 = lane_noise_scale(cfg, dt1, os);
  double Ym[N * D], YL[N * N];
  lane_revert_prior<N, D>(Q, P, PI, s1, Pm, PL, Ym, YL, w.cond(to));
  if (cfg.strategy == PDQ_FIXEDPOINT) lane_merge<N, D>(true, w.cond(to), w.cond(to), w.cond(to));
  const CondRef<N, D> ext = w.cond(from);
  lane_write_checkpoint<N, D>(c, k, t1, Pm, PL, os, nsteps, &ext, false);
  // new interp_from: fixed point (ext_t.marginal, identity); fixed interval ext_t
  w.store_mL(from, Pm, PL);
  w.at(from, LS_T) = t1;
  w.at(from, LS_IDENT) = cfg.strategy == PDQ_FIXEDPOINT ? 1.0 : 0.0;
  w.at(to, LS_IDENT) = 0.0;
}

// userfriendly_output + strategy.finalize (solvers.py:439-480, 601-633,
// 735-767; estimators_and_losses.py:380-391, 437-470).
template <int N, int D>
__device__ __forceinline__ void lane_finalize(const LaneCtx<N, D>& c, int sf) {
  const pdq_config& cfg = c.a->cfg;
  const pdq_solution& o = c.a->out;
  const int K = c.a->K, Km1 = K - 1;
  const size_t inst = (size_t)c.inst;
  double scale = 1.0;
  const bool mle = cfg.solver == PDQ_SOLVER_MLE;
  if (mle) {
    scale = c.w.at(sf, LS_MLE_RUN);
    if (cfg.correct_asymptotic_underconfidence) {
      const double ns = (double)o.num_steps[inst * Km1 + Km1 - 1];
      scale = scale / sqrt(ns);
    }
    if (o.output_scale)
      for (int k = 0; k < Km1; ++k) o.output_scale[inst * K + k] = scale;
  } else if (cfg.solver == PDQ_SOLVER && o.output_scale) {
    for (int k = 0; k < Km1; ++k) o.output_scale[inst * K + k] = 1.0;
  }
  if (cfg.strategy == PDQ_FILTER) {
    if (mle && o.chol)
      for (size_t e = 0; e < (size_t)K * N * N; ++e) o.chol[inst * K * N * N + e] = scale * o.chol[inst * K * N * N + e];
    return;
  }
  double Sm[N * D], SLf[N * N];
  c.w.load_mL(sf, Sm, SLf);
  const CondRef<N, D> cond = c.w.cond(sf);
  const bool ident = c.w.at(sf, LS_IDENT) != 0.0;
  if (mle) {
    for (size_t e = 0; e < (size_t)K * N * N; ++e) o.filt_chol[inst * K * N * N + e] = scale * o.filt_chol[inst * K * N * N + e];
    for (size_t e = 0; e < (size_t)Km1 * N * N; ++e)
      o.cond_chol[inst * Km1 * N * N + e] = scale * o.cond_chol[inst * Km1 * N * N + e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) SLf[e] = scale * SLf[e];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) cond.setL(i, j, scale * cond.L(i, j));
  }
  // rv_at_t1 = posterior1.conditional.marginalise(posterior1.marginal)
  double Xm[N * D], XL[N * N];
  if (ident) {
#pragma unroll
    for (int e = 0; e < N * D; ++e) Xm[e] = Sm[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = SLf[e];
  } else {
    lane_marginalise_cond<N, D>(cond, Sm, SLf, Xm, XL);
  }
  const size_t rowK = inst * K + (K - 1);
  lane_write_normal<N, D>(Xm, XL, o.term_mean ? o.term_mean + inst * N * D : nullptr,
                          o.term_chol ? o.term_chol + inst * N * N : nullptr);
  lane_write_normal<N, D>(Xm, XL, o.mean ? o.mean + rowK * N * D : nullptr, o.chol ? o.chol + rowK * N * N : nullptr);
  // backward pass: x_j = cond_j.marginalise(x_{j+1}), j = K-2 .. 0
  for (int j = Km1 - 1; j >= 0; --j) {
    const size_t idx = inst * Km1 + j;
    LCond<N, D> cj;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k) {
        cj.A(i, k) = o.cond_A[idx * N * N + i * N + k];
        cj.setL(i, k, o.cond_chol[idx * N * N + i * N + k]);
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < D; ++col) cj.m(i, col) = o.cond_mean[idx * N * D + i * D + col];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cj.tl(i) = o.cond_tl[idx * N + i];
      cj.to(i) = o.cond_to[idx * N + i];
    }
    double Ym[N * D], YL[N * N];
    lane_marginalise_cond<N, D>(cj, Xm, XL, Ym, YL);
    const size_t row = inst * K + j;
    lane_write_normal<N, D>(Ym, YL, o.mean ? o.mean + row * N * D : nullptr, o.chol ? o.chol + row * N * N : nullptr);
#pragma unroll
    for (int e = 0; e < N * D; ++e) Xm[e] = Ym[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
  }
}

// The adaptive loop's state of one lane (solvers_via_adaptive_steps.py:
// 46-375): the three slots' roles, the checkpoint index, the controller.
struct LaneLoop {
  int sf, ifr, pr, k, att, status;
  bool second_half;
  double dt, cstate;
  long attempts, accepted;
};

// The work an iteration leaves for after its part 2:
// * on: the fixed-point merge cond(res) = cond(self) o cond(other), with
//   `ident` when cond(self) is the identity (lane_merge);
// * ckpt: checkpoint k from slot `self` (marginal and conditional, the
//   conditional an identity with `ident`), with the output scale and step
//   count of slot `other` and the time of slot `self` (own_t) or ts[k].
struct LaneMergeReq {
  int on, ident, self, other, res, ckpt, own_t, k;
};

// What the first part of an iteration hands to the second.
template <int N, int D>
struct LaneIterVals {
  int op, ident;
  int fin;  // the conditional lane_revert_finish completes: lane_fin_pack, or -1
  double t1, tsf, h, tn, sigma, err, h0, os;
  double P[N], PI[N], b[D], Xm[N * D], Pm[N * D], PL[N * N];
};

enum { LOP_STEP = 0, LOP_BEYOND1 = 1, LOP_BEYOND2 = 2, LOP_AT = 3 };

template <int N, int D>
__device__ __forceinline__ StageRef<N, D> lane_stage(const LaneCtx<N, D>& c, int slot) {
  return StageRef<N, D>{c.w.cond(slot).p, c.stage_tail, c.stage_stride};
}

// What lane_revert_finish completes: the conditional's slot, the slot holding
// the reverted state's mean and factor, and lane_revert_prior's rescaling.
__device__ __forceinline__ int lane_fin_pack(int bws, int src, bool rescale) {
  return bws | (src << 2) | ((rescale ? 1 : 0) << 4);
}

__device__ __forceinline__ LaneMergeReq lane_no_req() {
  LaneMergeReq r;
  r.on = r.ident = r.self = r.other = r.res = r.ckpt = r.own_t = r.k = 0;
  return r;
}

// A request packed into two ints (flags and 2-bit slot indices, then k), as
// the merging wavefront reads it from LDS; the serial schedule takes the
// same round trip.
__device__ __forceinline__ void lane_pack_req(const LaneMergeReq& r, int* dst) {
  dst[0] = (r.on & 1) | ((r.ident & 1) << 1) | ((r.self & 3) << 2) | ((r.other & 3) << 4) | ((r.res & 3) << 6) |
           ((r.ckpt & 1) << 8) | ((r.own_t & 1) << 9);
  dst[1] = r.k;
}
__device__ __forceinline__ LaneMergeReq lane_unpack_req(const int* src) {
  const int v = src[0];
  LaneMergeReq r;
  r.on = v & 1;
  r.ident = (v >> 1) & 1;
  r.self = (v >> 2) & 3;
  r.other = (v >> 4) & 3;
  r.res = (v >> 6) & 3;
  r.ckpt = (v >> 8) & 1;
  r.own_t = (v >> 9) & 1;
  r.k = src[1];
  return r;
}

// prior init (from_mean_and_std): mean = Taylor coefficients, chol =
// init_std I, in slot sf; interp_from = a copy; checkpoint 0.
template <int N, int D>
__device__ __forceinline__ void lane_init_instance(const LaneCtx<N, D>& c, const double* ts, int sf, int ifr) {
  const pdq_config& cfg = c.a->cfg;
  const LaneWs<N, D>& w = c.w;
  const double* tc = c.a->tcoeffs + (size_t)c.inst * N * D;
  double m[N * D], L[N * N];
#pragma unroll
  for (int e = 0; e < N * D; ++e) m[e] = tc[e];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) L[i * N + j] = i == j ? (cfg.init_std_rows ? cfg.init_std_rows[i] : cfg.init_std) : 0.0;
#pragma unroll
  for (int e = 0; e < SlotLayout<N, D>::size; ++e) w.at(sf, e) = 0.0;
  w.store_mL(sf, m, L);
  w.at(sf, LS_T) = ts[0];
  w.at(sf, LS_OS) = 1.0;
  w.at(sf, LS_NSTEPS) = 0.0;
  w.at(sf, LS_IDENT) = cfg.strategy != PDQ_FILTER ? 1.0 : 0.0;
  w.copy_slot(ifr, sf);
  lane_write_checkpoint<N, D, CondRef<N, D>>(c, 0, ts[0], m, L, 1.0, 0.0, nullptr, false);
}

// `valid`: inst names an instance of the launch (per-instance inputs are read).
__device__ __forceinline__ LaneLoop lane_loop_start(const pdq_config& cfg, int inst, bool valid = true) {
  LaneLoop s;
  s.sf = 0;
  s.ifr = 1;
  s.pr = 2;
  s.k = 1;
  s.att = 0;
  s.status = PDQ_STATUS_OK;
  s.second_half = false;
  s.dt = cfg.dt0_per_instance && valid ? cfg.dt0_per_instance[inst] : cfg.dt0;
  s.cstate = 1.0;
  s.attempts = 0;
  s.accepted = 0;
  return s;
}

// The operation of the next iteration of the adaptive loop (needs s.k < K).
template <int N, int D>
__device__ __forceinline__ int lane_next_op(const pdq_config& cfg, const double* ts, const LaneLoop& s,
                                            const LaneWs<N, D>& w) {
  const double t1 = ts[s.k];
  const double tsf = w.at(s.sf, LS_T);
  if (s.second_half) return LOP_BEYOND2;
  if (tsf + cfg.eps < t1) return LOP_STEP;
  if (tsf > t1 + cfg.eps) return LOP_BEYOND1;
  return LOP_AT;
}

// The adaptive loop as a state machine: every iteration performs one
// operation per lane -- a step attempt, one half of interpolate_fwd beyond
// t1, or interpolate_fwd_at_t1 -- in the reference's order.  Step attempts
// and both halves of the interpolation share one prediction (revert), so
// lanes of a wavefront that are at different points of their solves execute
// it together instead of one after the other.
//
// Part 1: choose the operation, calibrate (solver_dynamic steps), and
// predict: the filter marginalises the transition; the smoothers revert it
// into the conditional of slot `bws`.
//
// Part 2: the step's update, error estimate and control decision, or the
// interpolation's bookkeeping and checkpoint (interpolate_fwd_at_t1 is done
// here whole).  It returns the fixed-point smoother's merge of the new
// backward conditional behind the accumulated one as a request, to be run
// after part 2 and before the next iteration's part 1 or beside that part 1:
// * a step's merge, cond(proposed) = cond(step_from) o cond(proposed), is
//   wanted only if the step is accepted and the next operation is not the
//   first half of an interpolation (which replaces that conditional);
// * the first half reverts into the proposed slot (unused between steps);
//   its merge is cond(interp_from) = cond(interp_from) o cond(proposed);
// * the second half reverts into step_from's conditional, already merged
//   behind an identity (lane_revert_prior's rescaling), and asks for none.
// The requested merge reads and writes only conditionals that the next part
// 1 neither reads nor writes: after an accepted step the next step reverts
// into the new proposed slot, the old interp_from; the second half writes
// step_from's conditional while the first half's merge uses interp_from's
// and proposed's; interpolate_fwd_at_t1 does nothing in part 1.
template <int N, int D>
__device__ __forceinline__ void lane_adapt_part1(const LaneCtx<N, D>& c, const double* ts, LaneLoop& s,
                                                 LaneIterVals<N, D>& v) {
  const pdq_config& cfg = c.a->cfg;
  const LaneWs<N, D>& w = c.w;
  v.t1 = ts[s.k];
  v.tsf = w.at(s.sf, LS_T);
  v.fin = -1;
  v.op = lane_next_op<N, D>(cfg, ts, s, w);
  if (v.op == LOP_AT) return;
  // inputs of the shared prediction
  const int xs = v.op == LOP_STEP ? s.sf : s.ifr;
  double XL[N * N];
  w.load_mL(xs, v.Xm, XL);
  v.tn = 0.0;
  v.sigma = 1.0;
  v.err = 0.0;
  v.h0 = 0.0;
  if (v.op == LOP_STEP) {
    v.h = s.dt;
    if (cfg.clip_dt) v.h = fmin(v.h, v.t1 - v.tsf);
    v.tn = v.tsf + v.h;
  } else if (v.op == LOP_BEYOND1) {
    v.h = v.t1 - w.at(s.ifr, LS_T);  // interp_from -> t1
  } else {
    v.h = v.tsf - v.t1;  // t1 -> step_from
  }
  lane_precon<N>(v.h, v.P, v.PI);
  if (v.op == LOP_STEP && cfg.solver == PDQ_SOLVER_DYNAMIC)
    lane_calibrate<N, D>(c, v.Xm, v.P, v.PI, v.h, v.tn, v.b, v.h0, v.sigma, v.err, true);
  v.os = w.at(s.sf, LS_OS);
  const double scale = v.op == LOP_STEP ? v.sigma : v.os;
  // step: revert into proposed; first half: revert into proposed (fixed
  // point, merged behind interp_from's conditional in part 2's request) or
  // straight into interp_from's (fixed interval); second half: revert into
  // step_from's, for the fixed point smoother already merged behind an
  // identity
  const bool fp = cfg.strategy == PDQ_FIXEDPOINT;
  int bws;
  if (v.op == LOP_STEP) bws = s.pr;
  else if (v.op == LOP_BEYOND1) bws = fp ? s.pr : s.ifr;
  else bws = s.sf;
  v.ident = w.at(v.op == LOP_BEYOND1 ? s.ifr : s.sf, LS_IDENT) != 0.0;
  PDQ_PHASE_BOUNDARY();
  const double ns = lane_noise_scale(cfg, v.h, scale);
  if (cfg.strategy == PDQ_FILTER) {
    lane_marginalise_prior<N, D>(c.a->Q, v.P, v.PI, ns, v.Xm, XL, v.Pm, v.PL);
  } else {
    // the smoothers: the prediction here, the conditional in
    // lane_revert_finish (beside part 2); the first half's source mean,
    // interp_from's, is overwritten by part 2, so it goes with the stage into
    // the proposed slot (whose mean nothing reads until a step writes it)
    lane_revert_left<N, D>(c.a->Q, v.P, v.PI, ns, v.Xm, XL, v.Pm, v.PL, lane_stage(c, bws));
    int src = xs;
    if (v.op == LOP_BEYOND1) {
#pragma unroll
      for (int i = 0; i < N * D; ++i) w.at(s.pr, LaneWs<N, D>::SL::m + i) = v.Xm[i];
      src = s.pr;
    }
    v.fin = lane_fin_pack(bws, src, fp && v.op == LOP_BEYOND2);
  }
  PDQ_PHASE_BOUNDARY();
}

// The fixed-point smoother's backward conditional of an iteration's part 1
// (lane_revert_right from the stage part 1 left in the conditional's storage
// and the stage tail, and the reverted state's mean in slot `src`).  Part 2
// reads none of these, and writes neither the conditional, the tail nor that
// slot's mean: this may run beside it.
template <int N, int D>
__device__ __forceinline__ void lane_revert_finish(const LaneCtx<N, D>& c, int fin) {
  if (fin < 0) return;
  double m[N * D];
  const int src = (fin >> 2) & 3;
#pragma unroll
  for (int i = 0; i < N * D; ++i) m[i] = c.w.at(src, LaneWs<N, D>::SL::m + i);
  lane_revert_right<N, D>(m, lane_stage(c, fin & 3), c.w.cond(fin & 3), ((fin >> 4) & 1) != 0);
}

// The work part 2 left: the merge (ssm_impl_isotropic.py:91-105) or the
// checkpoint.
template <int N, int D>
__device__ __forceinline__ void lane_run_merge(const LaneCtx<N, D>& c, const double* ts, const LaneMergeReq& rq) {
  const LaneWs<N, D>& w = c.w;
  if (rq.on) {
    lane_merge<N, D>(rq.ident != 0, w.cond(rq.self), w.cond(rq.other), w.cond(rq.res));
  } else if (rq.ckpt) {
#ifdef PDQ_HOST_EMU
    if (rq.k < 0 || rq.k >= c.a->K || rq.self < 0 || rq.self > 2 || rq.other < 0 || rq.other > 2) __builtin_trap();
#endif
    double m[N * D], L[N * N];
    w.load_mL(rq.self, m, L);
    const CondRef<N, D> cond = w.cond(rq.self);
    const double t = rq.own_t ? w.at(rq.self, LS_T) : ts[rq.k];
    lane_write_checkpoint<N, D>(c, rq.k, t, m, L, w.at(rq.other, LS_OS), w.at(rq.other, LS_NSTEPS), &cond,
                                rq.ident != 0);
  }
}

template <int N, int D>
__device__ __forceinline__ LaneMergeReq lane_adapt_part2(const LaneCtx<N, D>& c, const double* ts, LaneLoop& s,
                                                         LaneIterVals<N, D>& v) {
  const pdq_config& cfg = c.a->cfg;
  const LaneWs<N, D>& w = c.w;
  const bool smoother = cfg.strategy != PDQ_FILTER;
  const bool fp = cfg.strategy == PDQ_FIXEDPOINT;
  const int K = c.a->K;
  LaneMergeReq rq = lane_no_req();
  if (v.op == LOP_AT) {
    // interpolate_fwd_at_t1 (solvers.py:271-315): checkpoint k is step_from
    // (written after part 2), interp_from becomes a copy of it
    rq.ckpt = 1;
    rq.own_t = 1;
    rq.k = s.k;
    rq.self = s.sf;
    rq.other = s.sf;
    rq.ident = w.at(s.sf, LS_IDENT) != 0.0;
    w.copy_slot(s.ifr, s.sf);
    if (fp) {
      w.at(s.sf, LS_IDENT) = 1.0;
      w.at(s.ifr, LS_IDENT) = 1.0;
    }
    ++s.k;
  } else if (v.op == LOP_STEP) {
    const double ep = lane_finish_step<N, D>(c, s.sf, s.pr, v.h, v.tn, v.P, v.PI, v.Xm, v.Pm, v.PL, v.b, v.h0,
                                             v.sigma, v.err, true);
    s.dt = lane_control(cfg, v.h, s.cstate, ep);
    ++s.att;
    ++s.attempts;
    if (!isfinite(ep) || !isfinite(s.dt)) {
      s.status = PDQ_STATUS_NONFINITE;
    } else if (ep < 1.0 && (s.dt == 0.0 || v.tsf + s.dt == v.tsf)) {
      s.status = PDQ_STATUS_DT_UNDERFLOW;
    } else if (ep >= 1.0 && w.at(s.pr, LS_T) == v.tsf) {
      s.status = PDQ_STATUS_DT_UNDERFLOW;  // accepted without advancing t
    } else if ((cfg.max_attempts > 0 && s.att >= cfg.max_attempts && ep < 1.0) || s.attempts >= kMaxTotalAttempts) {
      s.status = PDQ_STATUS_MAX_ATTEMPTS;
    } else if (ep >= 1.0) {
      ++s.accepted;
      s.att = 0;
      rq.ident = v.ident;
      rq.self = s.sf;
      rq.other = s.pr;
      rq.res = s.pr;
      const int old_if = s.ifr;  // interp_from <- step_from, step_from <- proposed
      s.ifr = s.sf;
      s.sf = s.pr;
      s.pr = old_if;
      rq.on = fp && lane_next_op<N, D>(cfg, ts, s, w) != LOP_BEYOND1;
    }
  } else if (v.op == LOP_BEYOND1) {
    const double nsteps = w.at(s.sf, LS_NSTEPS);
    if (!smoother) {
      lane_write_checkpoint<N, D, CondRef<N, D>>(c, s.k, v.t1, v.Pm, v.PL, v.os, nsteps, nullptr, false);
      w.store_mL(s.ifr, v.Pm, v.PL);
      w.at(s.ifr, LS_T) = v.t1;
      ++s.k;
    } else {
      w.store_mL(s.ifr, v.Pm, v.PL);  // ext_t's marginal: the new interp_from's
      s.second_half = true;
      rq.ident = v.ident;
      rq.self = s.ifr;
      rq.other = s.pr;
      rq.res = s.ifr;
      rq.on = fp;
    }
  } else {
    // interpolated = (ext_t.marginal, ext_t.conditional), the checkpoint
    // (written after part 2)
    rq.ckpt = 1;
    rq.own_t = 0;
    rq.k = s.k;
    rq.self = s.ifr;
    rq.other = s.sf;
    rq.ident = 0;
    w.at(s.ifr, LS_T) = v.t1;
    w.at(s.ifr, LS_IDENT) = fp ? 1.0 : 0.0;
    w.at(s.sf, LS_IDENT) = 0.0;
    s.second_half = false;
    ++s.k;
  }
  (void)K;
  return rq;
}

template <int N, int D>
__device__ __forceinline__ void lane_end_instance(const LaneCtx<N, D>& c, const LaneLoop& s) {
  if (s.status == PDQ_STATUS_OK) lane_finalize<N, D>(c, s.sf);
  const pdq_solution& o = c.a->out;
  const size_t inst = (size_t)c.inst;
  if (o.attempts) o.attempts[inst] = s.attempts;
  if (o.accepted) o.accepted[inst] = s.accepted;
  if (o.status) o.status[inst] = s.status;
}

template <int N, int D>
__device__ __forceinline__ LaneCtx<N, D> lane_ctx(const LArgs& a, int inst, double* slots) {
  LaneCtx<N, D> c;
  c.a = &a;
  c.inst = inst;
  c.prm = a.params + (size_t)inst * a.cfg.n_params;
  c.w.base = slots;
  c.lin_set = 0;
  c.stage_tail = nullptr;
  c.stage_stride = 1;
  return c;
}

// One instance: init, the save_at / fixed-grid loop, finalize, counters.
// `slots` is the lane's slot storage (its LDS column on the device).  The
// adaptive loop runs part 1, part 2 and the merge part 2 asks for, in turn.
template <int N, int D>
__device__ __forceinline__ void lane_solve_instance(const LArgs& a, int inst, double* slots) {
  LaneCtx<N, D> c = lane_ctx<N, D>(a, inst, slots);
  const int K = a.K;
  const double* ts = a.times + (size_t)inst * K;
  LaneLoop s = lane_loop_start(a.cfg, inst);
  lane_init_instance<N, D>(c, ts, s.sf, s.ifr);
  if (a.mode == LMODE_FIXED_GRID) {
    for (int k = 1; k < K; ++k) {
      c.lin_set = pdq_lin_set(a.cfg, k - 1);
      lane_step<N, D>(c, s.sf, s.pr, ts[k] - ts[k - 1], false);
      ++s.attempts;
      ++s.accepted;
      lane_write_slot<N, D>(c, k, s.pr);
      const int tmp = s.sf;
      s.sf = s.pr;
      s.pr = tmp;
    }
  } else {
    double tail[RevertStage<N, D>::tail > 0 ? RevertStage<N, D>::tail : 1];
    c.stage_tail = tail;
    while (s.k < K && s.status == PDQ_STATUS_OK) {
      LaneIterVals<N, D> v;
      c.lin_set = pdq_lin_set(a.cfg, s.accepted);
      lane_adapt_part1<N, D>(c, ts, s, v);
      lane_revert_finish<N, D>(c, v.fin);
      int packed[2];
      lane_pack_req(lane_adapt_part2<N, D>(c, ts, s, v), packed);
      lane_run_merge<N, D>(c, ts, lane_unpack_req(packed));
    }
  }
  lane_end_instance<N, D>(c, s);
}

// The smoothers' adaptive solves run on two wavefronts (lane_solve_split):
// the second finishes the reverts' conditionals and runs the merges and
// checkpoint writes; the filter runs on one.
constexpr bool lane_split_strategy(int strategy) {
  return strategy == PDQ_FIXEDPOINT || strategy == PDQ_FIXEDINTERVAL;
}

template <int STRATEGY>
constexpr int lane_block_threads() {
  return lane_split_strategy(STRATEGY) ? 3 * kLaneBlock : kLaneBlock;
}

// LDS bytes of one workgroup: three slots per lane; with the second
// wavefront, 64 requests (two ints each), the loop flag (in an eight-byte
// cell), 64 lane_revert_finish codes and the lanes' revert stage tails
// (element-major) after them.
template <int N, int D, int STRATEGY = PDQ_FILTER>
constexpr size_t lane_lds_bytes() {
  return sizeof(double) * kLaneBlock * 3 * SlotLayout<N, D>::size +
         (lane_split_strategy(STRATEGY) ? sizeof(double) * (kLaneBlock + 1) + sizeof(int) * kLaneBlock +
                                           sizeof(double) * kLaneBlock * RevertStage<N, D>::tail
                                     : 0);
}

#ifndef PDQ_HOST_EMU

// The smoothers' adaptive solve on two wavefronts: wavefront 0 runs parts 1
// and 2 of every iteration for the workgroup's 64 instances (one lane each,
// as lane_solve_instance); wavefront 1 runs the merge or checkpoint that
// part 2 of an iteration asks for while wavefront 0 runs part 1 of the next
// iteration, which touches none of the merge's conditionals (see
// lane_adapt_part2).  Lane j of both wavefronts works on instance j's LDS
// column only.  Per iteration: [part 1 (wave 0) | the previous request's
// merge (wave 1)] barrier [part 2, the new request and whether any lane goes
// on (wave 0)] barrier.  Both wavefronts read the posted flag after the same
// barrier, so they pass the same number of barriers and leave the loop
// together; a request and the flag are rewritten only after the barrier that
// follows their last read.  A request posted in the final iteration is not
// taken up by wavefront 1 (it leaves the loop); wavefront 0 runs it after
// the loop.
template <int N, int D>
__device__ __forceinline__ void lane_solve_split(const LArgs& a, double* lds) {
  const int l = threadIdx.x & (kLaneBlock - 1);
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x / kLaneBlock);  // wave-uniform
  double* const slots = lds + l;
  int* const req = reinterpret_cast<int*>(lds + kLaneBlock * 3 * SlotLayout<N, D>::size);
  volatile int* const flag = req + 2 * kLaneBlock;
  int* const fin = req + 2 * kLaneBlock + 2;
  double* const tail = reinterpret_cast<double*>(fin + kLaneBlock);
  const int inst = blockIdx.x * kLaneBlock + l;
  const bool mine = inst < a.n_inst;
  LaneCtx<N, D> c = lane_ctx<N, D>(a, mine ? inst : 0, slots);
  c.stage_tail = tail + l;
  c.stage_stride = kLaneBlock;
  if (role == 0) {
    const int K = a.K;
    const double* ts = a.times + (size_t)(mine ? inst : 0) * K;
    LaneLoop s = lane_loop_start(a.cfg, inst, mine);
    if (mine) lane_init_instance<N, D>(c, ts, s.sf, s.ifr);
    bool run = mine && s.k < K;
    req[2 * l] = 0;
    req[2 * l + 1] = 0;
    const int any0 = __any(run);
    if (l == 0) *flag = any0;
    __syncthreads();
    LaneMergeReq last = lane_no_req();
    while (__builtin_amdgcn_readfirstlane(*flag)) {
      LaneIterVals<N, D> v;
      c.lin_set = pdq_lin_set(a.cfg, s.accepted);
      v.fin = -1;
      if (run) lane_adapt_part1<N, D>(c, ts, s, v);
      fin[l] = v.fin;
      __syncthreads();
      last = lane_no_req();
      if (run) last = lane_adapt_part2<N, D>(c, ts, s, v);
      lane_pack_req(last, req + 2 * l);
      run = run && s.k < K && s.status == PDQ_STATUS_OK;
      const int any = __any(run);
      if (l == 0) *flag = any;
      __syncthreads();
    }
    // the merging wavefront has left: the request of the final iteration
    // (the last checkpoint of the lanes that ended in it) is run here
    if (mine) {
      int packed[2];
      lane_pack_req(last, packed);
      lane_run_merge<N, D>(c, ts, lane_unpack_req(packed));
      lane_end_instance<N, D>(c, s);
    }
  } else {
    // wavefront 1: the merges, then the conditionals; wavefront 2: the
    // checkpoint writes (a lane's request is one or the other)
    const double* ts = a.times + (size_t)(mine ? inst : 0) * a.K;
    __syncthreads();
    while (__builtin_amdgcn_readfirstlane(*flag)) {
      LaneMergeReq rq = lane_unpack_req(req + 2 * l);
      if (role == 1) rq.ckpt = 0;
      else rq.on = 0;
      lane_run_merge<N, D>(c, ts, rq);
      __syncthreads();
      if (role == 1) lane_revert_finish<N, D>(c, fin[l]);
      __syncthreads();
    }
  }
}

// One lane per instance.  Each lane's slots live in LDS, element-major across
// the workgroup's lanes, read and written by that lane only (and, for the
// smoothers' adaptive solves, by the same lane of the merging
// wavefront: lane_solve_split).  The strategy and the solver are
// compile-time constants of the instantiation: the kernel overwrites them in
// its own copy of the arguments, so every test of them folds and the other
// strategies' and solvers' code (with the registers it would hold) is not
// compiled in.  Without OPTS the linearisation, the error estimator's and
// the calibration's options take their defaults the same way (ts0,
// error_residual_std without re-linearisation or per-unit-step scaling, no
// re-linearisation after the calibration), and the launcher picks this
// instantiation only for configurations that hold them; with OPTS they are
// read at run time.  The smoothers' launches cover every instance with one
// lane of wavefront 0 (no grid stride); the filter's strides over instances.
template <int N, int D, int STRATEGY, int SOLVER, bool OPTS>
__global__ void __launch_bounds__(lane_block_threads<STRATEGY>()) lane_kernel(LArgs args) {
  extern __shared__ double lds[];
  static_assert(lane_lds_bytes<N, D, STRATEGY>() <= 160 * 1024, "a workgroup's LDS fits in one CU's 160 KB");
  LArgs a = args;
  a.cfg.strategy = STRATEGY;
  a.cfg.solver = SOLVER;
  if (!OPTS) {
    a.cfg.init_std_rows = nullptr;
    a.cfg.constraint = PDQ_TS0;
    a.cfg.error_estimator = PDQ_ERROR_RESIDUAL_STD;
    a.cfg.error_re_linearize = 0;
    a.cfg.error_per_unit_step = 0;
    a.cfg.re_linearize_after_calibration = 0;
  }
  if constexpr (lane_split_strategy(STRATEGY)) {
    if (a.mode != LMODE_FIXED_GRID) {
      lane_solve_split<N, D>(a, lds);
    } else if (threadIdx.x < kLaneBlock) {
      const int inst = blockIdx.x * kLaneBlock + threadIdx.x;
      if (inst < a.n_inst) lane_solve_instance<N, D>(a, inst, lds + threadIdx.x);
    }
  } else {
    double* const slots = lds + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int inst = blockIdx.x * blockDim.x + threadIdx.x; inst < a.n_inst; inst += stride)
      lane_solve_instance<N, D>(a, inst, slots);
  }
}
#endif

}  // namespace pdq
