// Block-diagonal factorisation (ssm_impl_blockdiag.py) on gfx950: one lane
// per block, one lane group per instance.
//
// A block-diagonal model is d independent (nu+1) x (nu+1) square-root
// filters, one per ODE dimension, that share the step: the vector field
// couples the blocks through the solution's zeroth coefficients, and the
// accept/reject decision and the step size come from one error norm over all
// blocks.  Every other operation is per block, and per block the algebra is
// that of the isotropic factorisation with one column (D = 1): the same
// marginalise / revert / merge / update restatements of pdq_lane.h, with the
// block's own output scale (per-dimension calibration,
// BlockDiagNormal.residual_whitened_rms_flat, ssm_impl_blockdiag.py:276-292).
//
// Mapping.  Lane l of a group of G lanes owns block l of the group's
// instance for the whole solve: its three solution slots (step_from,
// interp_from, proposed) are register arrays, moved between slots by
// compile-time copies.  Per-instance scalars (t, step counts, the smoother's
// identity flags) are computed redundantly by every lane of the group from
// group-uniform values, so that all lanes take the same decisions.  The
// group's lanes meet at two points per step attempt:
//  * the linearisation: each lane publishes its block's zeroth coefficient
//    in LDS, the group synchronises, and each lane evaluates its component
//    f_k of the vector field (and, for ts1, the Jacobian's diagonal entry);
//  * the error norm: a group sum of the blocks' scaled squared errors; every
//    lane ends with the bitwise same total.
// Groups of G <= 64 lanes (G a power of two >= d) share a wavefront and
// synchronise within it; d > 64 takes one workgroup of G = 64 * ceil(d / 64)
// lanes (<= 512) per instance and synchronises with workgroup barriers,
// which every lane of the workgroup reaches because its control flow is
// uniform.  Lanes beyond d pad the group: they take part in every
// synchronisation and contribute exact zeros to the sums, and write nothing.
//
// The loops follow _ivpsolve/solvers_via_adaptive_steps.py:46-375 and
// solvers.py:205-767 in the order of the lane-per-instance kernel
// (pdq_lane.h) and the oracle, with the same status exits.
#pragma once

#include "pdq_lane.h"

#ifdef PDQ_HOST_EMU
// The CPU build runs a group as G threads (tests/emu): these hooks are the
// harness's barrier and lane exchange.
void pdq_emu_group_barrier();
double pdq_emu_shfl_xor(double v, int mask, int width);
double pdq_emu_shfl(double v, int src, int width);
#endif

namespace pdq {

// Largest block-diagonal dimension (one lane per block, <= 512 lanes per
// instance).
constexpr int kBdMaxDim = 512;
constexpr int kBdRedStride = kBdMaxDim / 64;  // one group sum buffer: a partial per wavefront

// Arguments of the block-diagonal kernel.
struct BdArgs {
  pdq_config cfg;
  int n_inst, K, mode;     // mode: LMODE_SAVE_AT or LMODE_FIXED_GRID
  int G;                   // lanes per instance
  const double* params;    // n_inst x n_params
  const double* tcoeffs;   // n_inst x (nu+1) x d
  const double* times;     // n_inst x K
  const double* Q;         // (nu+1) x (nu+1) latent IWP noise factor
  pdq_solution out;
};

// One instance's lane group.
struct BdGroup {
  int G;       // lanes (blocks incl. padding)
  int lane;    // = block index
  bool multi;  // the group is the whole (multi-wavefront) workgroup
  double* y;   // 2 G doubles of LDS: the published zeroth coefficients, two buffers
  double* red; // multi: one partial sum per wavefront, two buffers of kBdRedStride
  // The buffer each of the next linearisation and group sum uses.  They
  // alternate, so a write never meets a read of the same buffer from the
  // previous call: between the two calls that share a buffer lies the other
  // call's barrier, which every lane reaches after its reads.  Every lane
  // makes the same calls, so the toggles stay uniform.
  mutable int ybuf = 0, rbuf = 0;

  __device__ __forceinline__ void sync() const {
#ifdef PDQ_HOST_EMU
    pdq_emu_group_barrier();
#else
    if (multi) {
      __syncthreads();
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#endif
  }
  __device__ __forceinline__ double shfl_xor(double v, int m, int w) const {
#ifdef PDQ_HOST_EMU
    return pdq_emu_shfl_xor(v, m, w);
#else
    return __shfl_xor(v, m, w);
#endif
  }
  __device__ __forceinline__ double shfl(double v, int src, int w) const {
#ifdef PDQ_HOST_EMU
    return pdq_emu_shfl(v, src, w);
#else
    return __shfl(v, src, w);
#endif
  }
  // Group sum; every lane returns the bitwise same total.  One wavefront: a
  // butterfly, then lane 0's total.  Several: each wavefront's butterfly
  // total, then the wavefronts' totals added in wavefront order.
  __device__ __forceinline__ double sum(double v) const {
    if (!multi) {
      for (int m = G / 2; m >= 1; m >>= 1) v += shfl_xor(v, m, G);
      return shfl(v, 0, G);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m, 64);
    v = shfl(v, 0, 64);
    const int w = lane / 64, nw = G / 64;
    double* const r = red + rbuf * kBdRedStride;
    rbuf ^= 1;
    if ((lane & 63) == 0) r[w] = v;
    sync();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += r[i];
    return s;
  }
};

// One block's solution slot: mean, lower factor (full N x N with structural
// zeros above the diagonal), output scale, solver_mle's running scale, and
// (smoothers) the backward conditional.
template <int N>
struct BdSlot {
  double m[N];
  double L[N * N];
  double os;
  double mle_run;
  LCond<N, 1> c;
};

// A view of a register conditional with CondRef's interface (accessors that
// write through a const view), for the pdq_lane.h algebra.
template <int N>
struct LCondView {
  using CL = CondLayout<N, 1>;
  LCond<N, 1>* p;
  __device__ __forceinline__ double& at(int e) const { return p->v[e]; }
  __device__ __forceinline__ double& A(int i, int j) const { return p->v[CL::A + i * N + j]; }
  __device__ __forceinline__ double& m(int i, int) const { return p->v[CL::m + i]; }
  __device__ __forceinline__ double& tl(int i) const { return p->v[CL::tl + i]; }
  __device__ __forceinline__ double& to(int i) const { return p->v[CL::to + i]; }
  __device__ __forceinline__ double L(int i, int j) const { return j <= i ? p->v[CL::L + tri(i, j)] : 0.0; }
  __device__ __forceinline__ void setL(int i, int j, double x) const {
    if (j <= i) p->v[CL::L + tri(i, j)] = x;
  }
};

template <int N>
__device__ __forceinline__ LCondView<N> cv(LCond<N, 1>& c) {
  return LCondView<N>{&c};
}

// Group-uniform scalars of a slot.
struct BdScal {
  double t;
  double nsteps;
  double mle_n;
  bool ident;  // smoothers: the conditional is the identity
};

template <int N, int STRATEGY>
struct BdCtx {
  const BdArgs* a;
  const double* prm;
  BdGroup g;
  int inst;
  bool act;  // lane < d (padding lanes write nothing)
  uint32_t lin_set;  // the step's probe set of the Monte Carlo Jacobian handlers (pdq_lin_set)
  // Constant denominators of the step and their reciprocals (correctly
  // rounded, so that lane_div returns the correctly rounded quotient, bitwise
  // the division's): sqrt(d) of the error norm, and Burgers' 2 dx and dx dx.
  double sqd, rsqd, bg_den1, bg_rden1, bg_den2, bg_rden2;
};

// Burgers' component k (the registry's formula, vf_component in
// pdq_device.h): the periodic neighbours without an integer modulo, the two
// quotients from the instance's reciprocals.  Bitwise the registry's value.
template <int N, int STRATEGY>
__device__ __forceinline__ double bd_burgers(const BdCtx<N, STRATEGY>& c, const double* y, int d, int k) {
  const double visc = c.prm[0];
  const int km = k == 0 ? d - 1 : k - 1, kp = k + 1 == d ? 0 : k + 1;
  const double um = y[km], u0 = y[k], up = y[kp];
  return -u0 * lane_div(up - um, c.bg_den1, c.bg_rden1) + visc * lane_div(up - 2.0 * u0 + um, c.bg_den2, c.bg_rden2);
}

// The linearisation at the block means Xm (ts0: BlockDiagOdeTs0.linearize,
// ssm_impl_blockdiag.py:123-144; ts1: BlockDiagResidual.linearize, :147-183,
// with the exact Jacobian diagonal of jacobian_materialize, jacobians.py:
// 108-113): b = -f_k(u) (ts0) or r_k(m) - [h0, 1] m with h0 = -J_f[k][k]
// (ts1).  With the Monte Carlo handlers (calculate_diagonal_along_d of
// jacobian_monte_carlo_rev and _fwd, jacobians.py:250-271, 171-192) h0 is minus
// the probe mean of (v^T J)_k v_k or (J v)_k v_k over the probe set c.lin_set.  Every lane of the group calls it (it
// synchronises twice).
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_linearize(const BdCtx<N, STRATEGY>& c, const double* Xm, double t, double& b,
                                             double& h0, uint32_t set_offset = 0) {
  const pdq_config& cfg = c.a->cfg;
  const int d = cfg.ode_dim, k = c.g.lane;
  double* const y = c.g.y + c.g.ybuf * c.g.G;  // the buffers alternate: no barrier after the reads
  c.g.ybuf ^= 1;
  if (c.act) y[k] = Xm[0];
  c.g.sync();
  double f = 0.0, jkk = 0.0;
  if (c.act) {
    f = cfg.vector_field == PDQ_VF_BURGERS ? bd_burgers(c, y, d, k) : vf_component(cfg.vector_field, d, y, c.prm, t, k);
    if (cfg.constraint == PDQ_TS1)
      jkk = cfg.jacobian != PDQ_JAC_EXACT
                ? vf_mc_diag(cfg.vector_field, d, y, c.prm, k, cfg.jac_seed, c.lin_set + set_offset, cfg.jac_probes,
                             cfg.jacobian == PDQ_JAC_MONTE_CARLO_FWD)
                : vf_jac(cfg.vector_field, d, y, c.prm, k, k);
  }
  if (cfg.constraint != PDQ_TS1) {
    h0 = 0.0;
    b = -f;
    return;
  }
  h0 = -jkk;
  const double r = Xm[1] - f;
  b = r - (h0 * Xm[0] + Xm[1]);
}

// Per block: the calibrated error std and the reference (derivative `row` of
// the previous and proposed means), as error_residual_std / error_state_std
// produce them (solvers.py:925-996, 1041-1098); then the error power of the
// whole instance from group sums (error_norm_scale_then_rms /
// _rms_then_scale, solvers.py:770-811, over d errors and d references).
template <int N, int STRATEGY>
__device__ __forceinline__ double bd_error_power(const BdCtx<N, STRATEGY>& c, double err, int row, const double* prev_m,
                                                 const double* prop_m, double dt, int nord) {
  const pdq_config& cfg = c.a->cfg;
  const double e = nord == 1 ? err * dt / 1.0 : err * pow(dt, (double)nord) / factorial_d(nord);
  double ref = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r)
    if (r == row) ref = fmax(fabs(prev_m[r]), fabs(prop_m[r]));
  const double d = (double)cfg.ode_dim;
  double norm;
  if (cfg.error_norm == PDQ_NORM_SCALE_THEN_RMS) {
    const double rr = e / (cfg.atol + cfg.rtol * fabs(ref));
    const double tot = c.g.sum(c.act ? rr * rr : 0.0);
    norm = lane_div(sqrt(tot), c.sqd, c.rsqd);  // sqrt(tot) / sqrt(d)
  } else {
    const double sa = c.g.sum(c.act ? e * e : 0.0);
    const double sr = c.g.sum(c.act ? ref * ref : 0.0);
    norm = (sqrt(sa) / sqrt(d)) / (cfg.atol + cfg.rtol * (sqrt(sr) / sqrt(d)));
  }
  return lane_root_inv<N>(norm);  // pow(norm, -1/N), as the lane kernel's error power takes it
}

// The transition's marginalise with the block-diagonal |to_latent| on the
// factor (ssm_impl_blockdiag.py:28-43).
template <int N>
__device__ __forceinline__ void bd_marginalise_prior(const double* __restrict__ Q, const double* P, const double* PI,
                                                     double s, const double* Xm, const double* XL, double* Ym,
                                                     double* YL) {
  double API[N];
#pragma unroll
  for (int k = 0; k < N; ++k) API[k] = fabs(PI[k]);
  double big[2 * N * N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int col = 0; col < N; ++col) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double a = lane_iwp_a<N>(col, k);
        if (a != 0.0) acc += a * (API[k] * XL[k * N + r]);
      }
      big[r * N + col] = acc;
      big[(N + r) * N + col] = s * lane_iwp_q<N>(Q, col, r);
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double a = lane_iwp_a<N>(i, k);
      if (a != 0.0) acc += a * (PI[k] * Xm[k]);
    }
    Ym[i] = P[i] * (acc + 0.0);
  }
  lane_qr<2 * N, N, TriStackRows<N>>(big);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) YL[i * N + j] = fabs(P[i]) * big[j * N + i];
}

// A backward conditional's marginalise (estimators_and_losses.py:156-178)
// with |to_latent| on the factor (ssm_impl_blockdiag.py:28-43).
template <int N, class Cond>
__device__ __forceinline__ void bd_marginalise_cond(const Cond& c, const double* Xm, const double* XL, double* Ym,
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
      for (int k = 0; k < N; ++k) acc += c.A(col, k) * (fabs(tl[k]) * XL[k * N + r]);
      big[r * N + col] = acc;
      big[(N + r) * N + col] = c.L(col, r);
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) acc += c.A(i, k) * (tl[k] * Xm[k]);
    Ym[i] = c.to(i) * (acc + c.m(i, 0));
  }
  lane_qr<2 * N, N, TriStackRows<N>>(big);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) YL[i * N + j] = fabs(c.to(i)) * big[j * N + i];
}

// ---------------------------------------------------------------- outputs
template <int N>
__device__ __forceinline__ void bd_write_normal(const double* m, const double* L, double* gm, double* gL) {
  if (gm)
#pragma unroll
    for (int i = 0; i < N; ++i) gm[i] = m[i];
  if (gL)
#pragma unroll
    for (int i = 0; i < N * N; ++i) gL[i] = L[i];
}

// Backward conditional j of block k (identity when `ident`).
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_write_cond(const BdCtx<N, STRATEGY>& c, bool ident, const LCond<N, 1>& s, int j) {
  const pdq_solution& o = c.a->out;
  const int d = c.a->cfg.ode_dim, k = c.g.lane;
  const size_t idx = ((size_t)c.inst * (c.a->K - 1) + j) * d + k;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      o.cond_A[idx * N * N + i * N + q] = ident ? (i == q ? 1.0 : 0.0) : s.A(i, q);
      o.cond_chol[idx * N * N + i * N + q] = ident ? 0.0 : s.L(i, q);
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    o.cond_mean[idx * N + i] = ident ? 0.0 : s.m(i, 0);
    o.cond_tl[idx * N + i] = ident ? 1.0 : s.tl(i);
    o.cond_to[idx * N + i] = ident ? 1.0 : s.to(i);
  }
}

// Checkpoint k: the filter's marginal, or the smoothers' filtering marginal
// (and, with `cond`, the backward conditional that leads to it).
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_write_checkpoint(const BdCtx<N, STRATEGY>& c, int k, double t, const double* m,
                                                    const double* L, double os, double nsteps,
                                                    const LCond<N, 1>* cond, bool cond_ident) {
  const pdq_solution& o = c.a->out;
  const pdq_config& cfg = c.a->cfg;
  const int K = c.a->K, d = cfg.ode_dim, blk = c.g.lane;
  const size_t row = (size_t)c.inst * K + k;
  if (blk == 0) {
    if (o.t) o.t[row] = t;
    if (o.num_steps && k > 0) o.num_steps[(size_t)c.inst * (K - 1) + k - 1] = (int64_t)nsteps;
  }
  if (!c.act) return;
  if (o.output_scale && cfg.solver == PDQ_SOLVER_DYNAMIC) o.output_scale[row * d + blk] = os;
  const size_t mo = (row * d + blk) * N, co = (row * d + blk) * N * N;
  if (STRATEGY != PDQ_FILTER) {
    bd_write_normal<N>(m, L, o.filt_mean ? o.filt_mean + mo : nullptr, o.filt_chol ? o.filt_chol + co : nullptr);
    if (k > 0 && cond) bd_write_cond<N, STRATEGY>(c, cond_ident, *cond, k - 1);
  } else {
    bd_write_normal<N>(m, L, o.mean ? o.mean + mo : nullptr, o.chol ? o.chol + co : nullptr);
  }
}

template <int N, int STRATEGY>
__device__ __forceinline__ void bd_write_slot(const BdCtx<N, STRATEGY>& c, int k, const BdSlot<N>& s,
                                              const BdScal& z) {
  bd_write_checkpoint<N, STRATEGY>(c, k, z.t, s.m, s.L, s.os, z.nsteps, &s.c, z.ident);
}

template <int N>
__device__ __forceinline__ void bd_copy(BdSlot<N>& dst, const BdSlot<N>& src) {
#pragma unroll
  for (int i = 0; i < N; ++i) dst.m[i] = src.m[i];
#pragma unroll
  for (int i = 0; i < N * N; ++i) dst.L[i] = src.L[i];
  dst.os = src.os;
  dst.mle_run = src.mle_run;
#pragma unroll
  for (int e = 0; e < CondLayout<N, 1>::size; ++e) dst.c.v[e] = src.c.v[e];
}

// One step attempt from `sf` into `pr`: solver_dynamic.step
// (solvers.py:549-599), solver.step (:702-733) or solver_mle.step
// (:395-437), with the error estimate.  Returns error_power (with
// `estimate`).  Every lane of the group calls it.
template <int N, int STRATEGY>
__device__ __forceinline__ double bd_step(const BdCtx<N, STRATEGY>& c, const BdSlot<N>& sf, const BdScal& zs,
                                          BdSlot<N>& pr, BdScal& zp, double h, bool estimate) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const bool state_std = cfg.error_estimator == PDQ_ERROR_STATE_STD;
  const double tn = zs.t + h;
  double P[N], PI[N];
  lane_precon<N>(h, P, PI);
  const double s1 = lane_noise_scale(cfg, h, 1.0);
  double b = 0.0, h0 = 0.0, sigma = 1.0, err = 0.0;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    // (1) calibration at the output-scale-1 extrapolation
    double Um[N], UL[N * N];
    lane_apply_flat_mean<N, 1>(P, PI, sf.m, Um);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) UL[i * N + j] = fabs(P[i]) * (s1 * lane_iwp_q<N>(Q, i, j));
    bd_linearize<N, STRATEGY>(c, Um, tn, b, h0);
    double om, oL;
    lane_marginalise_fx<N, 1>(cfg, &b, h0, Um, UL, &om, oL);
    sigma = lane_rms_zero<1>(&om, oL);
    if (estimate)
      err = state_std ? lane_state_err<N, 1>(cfg, &b, h0, Um, UL, cfg.error_derivative_idx) : lane_residual_err(sigma, oL);
  }
  PDQ_PHASE_BOUNDARY();
  // (2) the prediction
  double Pm[N], PL[N * N];
  const double s = lane_noise_scale(cfg, h, sigma);
  if (STRATEGY == PDQ_FILTER) {
    bd_marginalise_prior<N>(Q, P, PI, s, sf.m, sf.L, Pm, PL);
  } else {
    lane_revert_prior<N, 1>(Q, P, PI, s, sf.m, sf.L, Pm, PL, cv<N>(pr.c));
    PDQ_PHASE_BOUNDARY();
    if (STRATEGY == PDQ_FIXEDPOINT) lane_merge<N, 1>(zs.ident, sf.c, cv<N>(pr.c), cv<N>(pr.c));
  }
  PDQ_PHASE_BOUNDARY();
  // (3) the update, the MLE running scale and the error estimate
  double Rm[N], RL[N * N], mo, oLu;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    double bu = b, hu = h0;
    if (cfg.re_linearize_after_calibration) bd_linearize<N, STRATEGY>(c, Pm, tn, bu, hu, 1);
    lane_update<N, 1>(cfg, &bu, hu, Pm, PL, Rm, RL, &mo, oLu);
    pr.mle_run = 0.0;
    zp.mle_n = 0.0;
  } else {
    bd_linearize<N, STRATEGY>(c, Pm, tn, b, h0);
    if (estimate) {
      double Um[N], UL[N * N], be = b, he = h0;
      lane_apply_flat_mean<N, 1>(P, PI, sf.m, Um);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) UL[i * N + j] = fabs(P[i]) * (s1 * lane_iwp_q<N>(Q, i, j));
      // re_linearize_before_error (solvers.py:951-956): a new linearisation
      // at the extrapolation
      if (cfg.error_re_linearize) bd_linearize<N, STRATEGY>(c, Um, tn, be, he);
      if (state_std) {
        err = lane_state_err<N, 1>(cfg, &be, he, Um, UL, cfg.error_derivative_idx);
      } else {
        double om, oLe;
        lane_marginalise_fx<N, 1>(cfg, &be, he, Um, UL, &om, oLe);
        err = lane_residual_err(lane_rms_zero<1>(&om, oLe), oLe);
      }
    }
    lane_update<N, 1>(cfg, &b, h0, Pm, PL, Rm, RL, &mo, oLu);
    if (cfg.solver == PDQ_SOLVER_MLE) {
      // bayes_rule_and_residual_whitened_rms_tree: observed = (mo, R_Y^T)
      const double nt = lane_rms_zero<1>(&mo, oLu);
      const double nd = zs.mle_n;
      const double x1 = sqrt(nd / (nd + 1)) * sf.mle_run;
      const double x2 = sqrt(1 / (nd + 1)) * nt;
      pr.mle_run = hypot(x1, x2);
      zp.mle_n = nd + 1;
      sigma = sf.os;
    } else {
      pr.mle_run = 0.0;
      zp.mle_n = 0.0;
      sigma = 1.0;
    }
  }
  double ep = 0.0;
  if (estimate) {
    const int err_row = state_std ? cfg.error_derivative_idx : 0;
    const int nord = (state_std ? cfg.error_derivative_idx : 1) + (cfg.error_per_unit_step ? 1 : 0);
    ep = bd_error_power<N, STRATEGY>(c, err, err_row, sf.m, Rm, h, nord);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) pr.m[i] = Rm[i];
#pragma unroll
  for (int i = 0; i < N * N; ++i) pr.L[i] = RL[i];
  pr.os = sigma;
  zp.t = tn;
  zp.nsteps = zs.nsteps + 1.0;
  zp.ident = false;
  return ep;
}

// ProbabilisticSolver.interpolate_fwd (solvers.py:205-269) from interp_from
// `fr` to step_from `sf` at t1: writes checkpoint k, updates both slots.
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_interp_beyond(const BdCtx<N, STRATEGY>& c, int k, double t1, BdSlot<N>& fr,
                                                 BdScal& zf, BdSlot<N>& sf, BdScal& zs) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const double os = sf.os;
  const double dt0 = t1 - zf.t, dt1 = zs.t - t1;
  double P[N], PI[N], Pm[N], PL[N * N];
  lane_precon<N>(dt0, P, PI);
  const double s0 = lane_noise_scale(cfg, dt0, os);
  if (STRATEGY == PDQ_FILTER) {
    bd_marginalise_prior<N>(Q, P, PI, s0, fr.m, fr.L, Pm, PL);
    bd_write_checkpoint<N, STRATEGY>(c, k, t1, Pm, PL, os, zs.nsteps, nullptr, false);
  } else {
    // ext_t = (Pm, PL) with the conditional interp_from.cond o bw (fixed
    // point; bw staged in step_from's conditional, rewritten below) or bw
    // (fixed interval), kept in interp_from's conditional
    if (STRATEGY == PDQ_FIXEDPOINT) {
      lane_revert_prior<N, 1>(Q, P, PI, s0, fr.m, fr.L, Pm, PL, cv<N>(sf.c));
      PDQ_PHASE_BOUNDARY();
      lane_merge<N, 1>(zf.ident, cv<N>(fr.c), cv<N>(sf.c), cv<N>(fr.c));
    } else {
      lane_revert_prior<N, 1>(Q, P, PI, s0, fr.m, fr.L, Pm, PL, cv<N>(fr.c));
    }
    PDQ_PHASE_BOUNDARY();
    // t -> t1 from ext_t's marginal: only the backward conditional is kept,
    // as step_from's (fixed point: identity o bw)
    lane_precon<N>(dt1, P, PI);
    const double s1 = lane_noise_scale(cfg, dt1, os);
    double Ym[N], YL[N * N];
    lane_revert_prior<N, 1>(Q, P, PI, s1, Pm, PL, Ym, YL, cv<N>(sf.c));
    if (STRATEGY == PDQ_FIXEDPOINT) lane_merge<N, 1>(true, cv<N>(sf.c), cv<N>(sf.c), cv<N>(sf.c));
    bd_write_checkpoint<N, STRATEGY>(c, k, t1, Pm, PL, os, zs.nsteps, &fr.c, false);
    zf.ident = STRATEGY == PDQ_FIXEDPOINT;
    zs.ident = false;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) fr.m[i] = Pm[i];
#pragma unroll
  for (int i = 0; i < N * N; ++i) fr.L[i] = PL[i];
  zf.t = t1;
}

// userfriendly_output + strategy.finalize (solvers.py:439-480, 601-633,
// 735-767; estimators_and_losses.py:380-391, 437-470), per block.
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_finalize(const BdCtx<N, STRATEGY>& c, BdSlot<N>& sf, const BdScal& zs) {
  const pdq_config& cfg = c.a->cfg;
  const pdq_solution& o = c.a->out;
  const int K = c.a->K, Km1 = K - 1, d = cfg.ode_dim, blk = c.g.lane;
  const size_t inst = (size_t)c.inst;
  if (!c.act) return;
  double scale = 1.0;
  const bool mle = cfg.solver == PDQ_SOLVER_MLE;
  if (mle) {
    scale = sf.mle_run;
    if (cfg.correct_asymptotic_underconfidence) scale = scale / sqrt(zs.nsteps);
    if (o.output_scale)
      for (int k = 0; k < Km1; ++k) o.output_scale[(inst * K + k) * d + blk] = scale;
  } else if (cfg.solver == PDQ_SOLVER && o.output_scale) {
    for (int k = 0; k < Km1; ++k) o.output_scale[(inst * K + k) * d + blk] = 1.0;
  }
  if (STRATEGY == PDQ_FILTER) {
    if (mle && o.chol)
      for (int k = 0; k < K; ++k)
        for (int e = 0; e < N * N; ++e) o.chol[((inst * K + k) * d + blk) * N * N + e] *= scale;
    return;
  }
  if (mle) {
    for (int k = 0; k < K; ++k)
      for (int e = 0; e < N * N; ++e) o.filt_chol[((inst * K + k) * d + blk) * N * N + e] *= scale;
    for (int k = 0; k < Km1; ++k)
      for (int e = 0; e < N * N; ++e) o.cond_chol[((inst * Km1 + k) * d + blk) * N * N + e] *= scale;
#pragma unroll
    for (int e = 0; e < N * N; ++e) sf.L[e] = scale * sf.L[e];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) sf.c.setL(i, j, scale * sf.c.L(i, j));
  }
  // rv_at_t1 = posterior1.conditional.marginalise(posterior1.marginal)
  double Xm[N], XL[N * N];
  if (zs.ident) {
#pragma unroll
    for (int e = 0; e < N; ++e) Xm[e] = sf.m[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = sf.L[e];
  } else {
    bd_marginalise_cond<N>(sf.c, sf.m, sf.L, Xm, XL);
  }
  const size_t rowK = inst * K + (K - 1);
  bd_write_normal<N>(Xm, XL, o.term_mean ? o.term_mean + (inst * d + blk) * N : nullptr,
                     o.term_chol ? o.term_chol + (inst * d + blk) * N * N : nullptr);
  bd_write_normal<N>(Xm, XL, o.mean ? o.mean + (rowK * d + blk) * N : nullptr,
                     o.chol ? o.chol + (rowK * d + blk) * N * N : nullptr);
  // backward pass: x_j = cond_j.marginalise(x_{j+1}), j = K-2 .. 0
  for (int j = Km1 - 1; j >= 0; --j) {
    const size_t idx = (inst * Km1 + j) * d + blk;
    LCond<N, 1> cj;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int q = 0; q < N; ++q) {
        cj.A(i, q) = o.cond_A[idx * N * N + i * N + q];
        cj.setL(i, q, o.cond_chol[idx * N * N + i * N + q]);
      }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cj.m(i, 0) = o.cond_mean[idx * N + i];
      cj.tl(i) = o.cond_tl[idx * N + i];
      cj.to(i) = o.cond_to[idx * N + i];
    }
    double Ym[N], YL[N * N];
    bd_marginalise_cond<N>(cj, Xm, XL, Ym, YL);
    const size_t row = inst * K + j;
    bd_write_normal<N>(Ym, YL, o.mean ? o.mean + (row * d + blk) * N : nullptr,
                       o.chol ? o.chol + (row * d + blk) * N * N : nullptr);
#pragma unroll
    for (int e = 0; e < N; ++e) Xm[e] = Ym[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
  }
}

// One instance on its lane group: init, the save_at / fixed-grid loop,
// finalize, counters.  Every lane of the group runs it with the same
// control flow.
template <int N, int STRATEGY>
__device__ __forceinline__ void bd_solve_instance(const BdArgs& a, const BdGroup& g, int inst) {
  const pdq_config& cfg = a.cfg;
  const int d = cfg.ode_dim, K = a.K, blk = g.lane;
  BdCtx<N, STRATEGY> c;
  c.a = &a;
  c.prm = a.params + (size_t)inst * cfg.n_params;
  c.g = g;
  c.inst = inst;
  c.act = blk < d;
  c.lin_set = 0;
  c.sqd = sqrt((double)d);
  c.rsqd = 1.0 / c.sqd;
  if (cfg.vector_field == PDQ_VF_BURGERS) {
    const double dx = c.prm[1];
    c.bg_den1 = 2.0 * dx;
    c.bg_rden1 = 1.0 / c.bg_den1;
    c.bg_den2 = dx * dx;
    c.bg_rden2 = 1.0 / c.bg_den2;
  } else {
    c.bg_den1 = c.bg_rden1 = c.bg_den2 = c.bg_rden2 = 1.0;
  }
  const double* ts = a.times + (size_t)inst * K;
  BdSlot<N> sf, fr, pr;
  BdScal zs, zf, zp;
  // prior init (BlockDiagNormal.from_mean_and_std, ssm_impl_blockdiag.py:
  // 235-253): block k's mean is the k-th entry of every Taylor coefficient;
  // the factor is init_std I
  const double* tc = a.tcoeffs + (size_t)inst * N * d;
#pragma unroll
  for (int i = 0; i < N; ++i) sf.m[i] = c.act ? tc[i * d + blk] : 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) sf.L[i * N + j] = i == j ? cfg.init_std : 0.0;
  sf.os = 1.0;
  sf.mle_run = 0.0;
#pragma unroll
  for (int e = 0; e < CondLayout<N, 1>::size; ++e) sf.c.v[e] = 0.0;
  zs.t = ts[0];
  zs.nsteps = 0.0;
  zs.mle_n = 0.0;
  zs.ident = STRATEGY != PDQ_FILTER;
  bd_copy<N>(fr, sf);
  bd_copy<N>(pr, sf);
  zf = zs;
  zp = zs;
  bd_write_checkpoint<N, STRATEGY>(c, 0, ts[0], sf.m, sf.L, 1.0, 0.0, nullptr, false);
  long attempts = 0, accepted = 0;
  int status = PDQ_STATUS_OK;
  if (a.mode == LMODE_FIXED_GRID) {
    for (int k = 1; k < K; ++k) {
      c.lin_set = pdq_lin_set(cfg, k - 1);
      bd_step<N, STRATEGY>(c, sf, zs, pr, zp, ts[k] - ts[k - 1], false);
      ++attempts;
      ++accepted;
      bd_write_slot<N, STRATEGY>(c, k, pr, zp);
      bd_copy<N>(sf, pr);
      zs = zp;
    }
  } else {
    double dt = cfg.dt0_per_instance ? cfg.dt0_per_instance[inst] : cfg.dt0, cstate = 1.0;
    int k = 1, att = 0;
    while (k < K && status == PDQ_STATUS_OK) {
      const double t1 = ts[k];
      if (zs.t + cfg.eps < t1) {
        // one step attempt (RejectionLoop.step_attempt)
        double h = dt;
        if (cfg.clip_dt) h = fmin(h, t1 - zs.t);
        c.lin_set = pdq_lin_set(cfg, accepted);
        const double ep = bd_step<N, STRATEGY>(c, sf, zs, pr, zp, h, true);
        dt = lane_control(cfg, h, cstate, ep);
        ++att;
        ++attempts;
        if (!isfinite(ep) || !isfinite(dt)) {
          status = PDQ_STATUS_NONFINITE;
        } else if (ep < 1.0 && (dt == 0.0 || zs.t + dt == zs.t)) {
          status = PDQ_STATUS_DT_UNDERFLOW;
        } else if (ep >= 1.0 && zp.t == zs.t) {
          status = PDQ_STATUS_DT_UNDERFLOW;  // accepted without advancing t
        } else if ((cfg.max_attempts > 0 && att >= cfg.max_attempts && ep < 1.0) || attempts >= kMaxTotalAttempts) {
          status = PDQ_STATUS_MAX_ATTEMPTS;
        } else if (ep >= 1.0) {
          ++accepted;
          att = 0;
          bd_copy<N>(fr, sf);  // interp_from <- step_from, step_from <- proposed
          zf = zs;
          bd_copy<N>(sf, pr);
          zs = zp;
        }
      } else if (zs.t > t1 + cfg.eps) {
        bd_interp_beyond<N, STRATEGY>(c, k, t1, fr, zf, sf, zs);
        ++k;
      } else {
        // interpolate_fwd_at_t1 (solvers.py:271-315)
        bd_write_slot<N, STRATEGY>(c, k, sf, zs);
        bd_copy<N>(fr, sf);
        zf = zs;
        if (STRATEGY == PDQ_FIXEDPOINT) {
          zs.ident = true;
          zf.ident = true;
        }
        ++k;
      }
    }
  }
  if (status == PDQ_STATUS_OK) bd_finalize<N, STRATEGY>(c, sf, zs);
  if (blk == 0) {
    const pdq_solution& o = a.out;
    if (o.attempts) o.attempts[inst] = attempts;
    if (o.accepted) o.accepted[inst] = accepted;
    if (o.status) o.status[inst] = status;
  }
}

#ifndef PDQ_HOST_EMU
// Groups of G <= 64 lanes, 64 / G instances per wavefront, grid-stride over
// instances.  A group's instances and its exits depend only on values the
// whole group shares, so each group leaves the loop as a unit.
template <int N, int STRATEGY>
__global__ void __launch_bounds__(64) bd_wave_kernel(BdArgs args) {
  __shared__ double y[2 * 64];
  BdArgs a = args;
  a.cfg.strategy = STRATEGY;
  const int G = a.G;
  const int gpw = 64 / G;
  BdGroup g;
  g.G = G;
  g.lane = threadIdx.x % G;
  g.multi = false;
  g.y = y + (threadIdx.x / G) * 2 * G;
  g.red = nullptr;
  const int stride = gridDim.x * gpw;
  for (int inst = blockIdx.x * gpw + threadIdx.x / G; inst < a.n_inst; inst += stride)
    bd_solve_instance<N, STRATEGY>(a, g, inst);
}

// One workgroup of G = 64 * ceil(d / 64) <= 512 lanes per instance, grid-
// stride over instances (uniform across the workgroup).
template <int N, int STRATEGY>
__global__ void __launch_bounds__(512) bd_block_kernel(BdArgs args) {
  __shared__ double y[2 * kBdMaxDim];
  __shared__ double red[2 * kBdRedStride];
  BdArgs a = args;
  a.cfg.strategy = STRATEGY;
  BdGroup g;
  g.G = a.G;
  g.lane = threadIdx.x;
  g.multi = true;
  g.y = y;
  g.red = red;
  for (int inst = blockIdx.x; inst < a.n_inst; inst += gridDim.x) bd_solve_instance<N, STRATEGY>(a, g, inst);
}
#endif

// Lanes per instance for dimension d (the launcher and the CPU build agree).
inline int bd_group_lanes(int d) {
  if (d > 64) return ((d + 63) / 64) * 64;
  int G = 1;
  while (G < d) G <<= 1;
  return G;
}

int launch_bd(BdArgs& a, hipStream_t st);

}  // namespace pdq
