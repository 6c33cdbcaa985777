// Lane-per-instance marginal log-likelihood and its gradient (isotropic).
//
// loss_lml_timeseries (estimators_and_losses.py:53-105) on the fixed-point
// smoother's Markov sequence: a forward pass on a fixed grid with a reset
// (interpolate_fwd_at_t1, :534-547) every `every` grid points, then
// MarkovSequence.evaluate_lml's reverse scan (:180-218) with the
// observation model to_derivative(0, std) and lstsq gains.
//
// The gradient is forward mode.  With ts0 (and the calibration under
// stop_gradient, solvers.py:568-575) no square-root factor depends on the
// parameters or the initial value, so every factor is shared by the primal
// and its tangents; only the mean recursions and the log-densities carry
// derivatives.  An instance runs on G lanes that each compute the shared
// factors and the primal mean (redundantly: they need no data from each
// other) and TPL of the n_tan = n_params + d tangent columns (lane g owns
// tangents g TPL .. g TPL + TPL - 1), so a lane's mean has
// MC = D (1 + TPL) columns: the primal, then its tangents, each D wide.  The
// device library instantiates TPL = 1 (G = n_tan lanes per instance) and
// TPL = 0 (the value alone, one lane); the code is generic in TPL.
//
// The per-step algebra is the lane solve's (pdq_lane.h) instantiated with MC
// mean columns.  Each lane keeps two solution slots in LDS (step_from and
// proposed; a fixed grid needs no interp_from) and writes the backward
// conditional of every reset interval to a scratch buffer in HBM,
// element-major across all lanes, which the reverse scan reads back.
#pragma once

#include "pdq_lane.h"

namespace pdq {

// A backward conditional stored element-major in the HBM scratch buffer:
// element e at p[e * stride] (stride = the number of lanes of the launch).
template <int N, int D>
struct CondStrided {
  using CL = CondLayout<N, D>;
  double* p;
  size_t stride;
  __device__ __forceinline__ double& at(int e) const { return p[(size_t)e * stride]; }
  __device__ __forceinline__ double A(int i, int j) const { return at(CL::A + i * N + j); }
  __device__ __forceinline__ double m(int i, int c) const { return at(CL::m + i * D + c); }
  __device__ __forceinline__ double tl(int i) const { return at(CL::tl + i); }
  __device__ __forceinline__ double to(int i) const { return at(CL::to + i); }
  __device__ __forceinline__ double L(int i, int j) const { return j <= i ? at(CL::L + tri(i, j)) : 0.0; }
};

// The observation block of to_derivative(0, std) (ssm_impl_isotropic.py:
// 217-236 via bayes_rule_and_logpdf_tree): [[std, 0], [(H L)^T, L^T]] with
// H = e_0.  Column 0 has non-zeros in rows 0 and 1 only (L is lower
// triangular); after that reflection the block is upper triangular.
struct ObsRows0 {
  static constexpr bool nz(int i, int r) { return i == 0 && r == 1; }
};

// ts0 linearisation with tangents (ssm_impl_isotropic.py:298-317 and its
// forward-mode derivative): b = -f(u); the tangent column of direction q is
// -(J_f(u) du_q + df/dtheta_q) (the last term for parameter directions only).
template <int N, int D, int TPL>
__device__ __forceinline__ void lane_lml_linearize(const pdq_config& cfg, const double* __restrict__ prm, int q0,
                                                   const double* Xm, double t, double* b) {
  constexpr int MC = D * (1 + TPL);
  double y[D];
#pragma unroll
  for (int j = 0; j < D; ++j) y[j] = Xm[j];
#pragma unroll
  for (int j = 0; j < D; ++j) b[j] = -vf_component(cfg.vector_field, D, y, prm, t, j);
#pragma unroll
  for (int tt = 0; tt < TPL; ++tt) {
    const int q = q0 + tt;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int jj = 0; jj < D; ++jj) acc += vf_jac(cfg.vector_field, D, y, prm, j, jj) * Xm[(1 + tt) * D + jj];
      if (q < cfg.n_params) acc += vf_jac_param(cfg.vector_field, D, y, prm, j, q);
      b[(1 + tt) * D + j] = -acc;
    }
  }
  (void)MC;
}

// What the prediction part of a fixed-grid step hands to its update part.
template <int N, int MC>
struct LmlStepVals {
  double P[N], PI[N], b[MC], Pm[N * MC], PL[N * N], sigma, tn;
};

// One fixed-grid step of the fixed-point smoother (solver.step /
// solver_dynamic.step, solvers.py:549-599 and 702-733) from slot `from` to
// slot `to` with tangent columns, in two parts: (1) calibration and the
// revert of the transition into cond(to); (2) the update into slot `to`'s
// marginal.  Between them (or beside part 1 of the next step) the merge
// cond(to) = cond(from) o cond(to), which neither part reads or writes.
// With `defer` the revert stops after its left half (lane_revert_left): the
// conditional of slot `to` is finished by lane_revert_right from the stage
// left in its storage and `stage_tail` (lane_lml_forward_split's second
// wavefront).
template <int N, int D, int TPL>
__device__ __forceinline__ void lane_lml_step_part1(const LaneCtx<N, D * (1 + TPL)>& c, int q0, int from, int to,
                                                    double h, LmlStepVals<N, D * (1 + TPL)>& v, bool defer = false,
                                                    double* stage_tail = nullptr) {
  constexpr int MC = D * (1 + TPL);
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const LaneWs<N, MC>& w = c.w;
  double Sm[N * MC], SLf[N * N];
  w.load_mL(from, Sm, SLf);
  v.tn = w.at(from, LS_T) + h;
  lane_precon<N>(h, v.P, v.PI);
  v.sigma = 1.0;
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    // calibration at the output-scale-1 extrapolation; sigma is a constant
    // for the tangents (stop_gradient)
    const double s1 = lane_noise_scale(cfg, h, 1.0);
    double Um[N * MC], col[N + 1];
    lane_apply_flat_mean<N, MC>(v.P, v.PI, Sm, Um);
    lane_lml_linearize<N, D, TPL>(cfg, c.prm, q0, Um, v.tn, v.b);
#pragma unroll
    for (int r = 0; r < N; ++r) col[r] = fabs(v.P[1]) * (s1 * lane_iwp_q<N>(Q, 1, r));
    col[N] = cfg.damp;
    double om[D];
#pragma unroll
    for (int k = 0; k < D; ++k) om[k] = Um[1 * MC + k] + v.b[k];
    lane_qr<N + 1, 1, FxColumnRows<N>>(col);
    v.sigma = lane_rms_zero<D>(om, col[0]);
  }
  PDQ_PHASE_BOUNDARY();
  if (defer)
    lane_revert_left<N, MC>(Q, v.P, v.PI, lane_noise_scale(cfg, h, v.sigma), Sm, SLf, v.Pm, v.PL,
                            StageRef<N, MC>{w.cond(to).p, stage_tail, kLaneBlock});
  else
    lane_revert_prior<N, MC>(Q, v.P, v.PI, lane_noise_scale(cfg, h, v.sigma), Sm, SLf, v.Pm, v.PL, w.cond(to));
  PDQ_PHASE_BOUNDARY();
}

template <int N, int D, int TPL>
__device__ __forceinline__ void lane_lml_step_part2(const LaneCtx<N, D * (1 + TPL)>& c, int q0, int from, int to,
                                                    LmlStepVals<N, D * (1 + TPL)>& v) {
  constexpr int MC = D * (1 + TPL);
  const pdq_config& cfg = c.a->cfg;
  const LaneWs<N, MC>& w = c.w;
  if (cfg.solver != PDQ_SOLVER_DYNAMIC || cfg.re_linearize_after_calibration)
    lane_lml_linearize<N, D, TPL>(cfg, c.prm, q0, v.Pm, v.tn, v.b);
  double Rm[N * MC], RL[N * N], mo[MC], oLu;
  lane_update<N, MC>(cfg, v.b, 0.0, v.Pm, v.PL, Rm, RL, mo, oLu);
  w.store_mL(to, Rm, RL);
  w.at(to, LS_T) = v.tn;
  w.at(to, LS_OS) = v.sigma;
  w.at(to, LS_NSTEPS) = w.at(from, LS_NSTEPS) + 1.0;
  w.at(to, LS_IDENT) = 0.0;
}

template <int N, int D, int TPL>
__device__ __forceinline__ void lane_lml_step(const LaneCtx<N, D * (1 + TPL)>& c, int q0, int from, int to,
                                              double h) {
  constexpr int MC = D * (1 + TPL);
  const LaneWs<N, MC>& w = c.w;
  LmlStepVals<N, MC> v;
  lane_lml_step_part1<N, D, TPL>(c, q0, from, to, h, v);
  lane_merge<N, MC>(w.at(from, LS_IDENT) != 0.0, w.cond(from), w.cond(to), w.cond(to));
  PDQ_PHASE_BOUNDARY();
  lane_lml_step_part2<N, D, TPL>(c, q0, from, to, v);
}

#ifndef PDQ_HOST_EMU
// The forward pass on two wavefronts (the LML kernel's SPLIT form): wavefront
// 0 runs the steps' two parts for 64 lanes, wavefront 1 runs each step's
// merge, and at an observation node the copy of the merged conditional to
// HBM, beside wavefront 0's next prediction part.  Three slots rotate (the
// step from `sf` reverts into `pr`; then sf <- pr, pr <- spare, spare <- sf),
// so the next prediction reads only the marginal of the merge's result slot
// and reverts into the third slot.  When the revert's stage fits in a
// conditional's storage, or the caller gives the lane's stage tail (stride
// 64) for the rest, wavefront 1 also finishes each step's backward
// conditional (lane_revert_right) beside wavefront 0's update: the update
// writes neither that conditional nor step_from's mean, which it reads.
// Wavefront 1 follows the slot rotation itself.  Every lane takes the same K - 1 steps, so
// both wavefronts pass 2 (K - 1) + 1 barriers; the last step's request is run
// by wavefront 0 after the loop.  Returns step_from's slot.  With ckpt_every
// > 0 wavefront 0 also stores step_from's whole slot before every
// ckpt_every-th step (checkpoint j, before step j ckpt_every, at ckpt + j |slot|
// n_lanes, element-major with stride n_lanes), once the previous step's merge
// has completed.
template <int N, int D, int TPL>
__device__ __forceinline__ int lane_lml_forward_split(const LaneCtx<N, D * (1 + TPL)>& c, const LmlArgs& a, int q0,
                                                      size_t glane, size_t n_lanes, bool active, int* req,
                                                      int ckpt_every = 0, double* ckpt = nullptr,
                                                      double* stage_tail = nullptr) {
  constexpr int MC = D * (1 + TPL);
  using CL = CondLayout<N, MC>;
  const bool defer = RevertStage<N, MC>::tail == 0 || stage_tail != nullptr;
  const LaneWs<N, MC>& w = c.w;
  const int K = a.K, every = a.every;
  const int l = threadIdx.x & (kLaneBlock - 1);
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x / kLaneBlock);
  // request: [0] = on | ident << 1 | self << 2 | res << 4, [1] = node index or -1
  auto run = [&](const int* r) {
    const int f = r[0];
    if (!(f & 1)) return;
    const int self = (f >> 2) & 3, res = (f >> 4) & 3;
    lane_merge<N, MC>(((f >> 1) & 1) != 0, w.cond(self), w.cond(res), w.cond(res));
    if (r[1] >= 0) {
      const CondStrided<N, MC> dst{a.scratch + (size_t)r[1] * CL::size * n_lanes + glane, n_lanes};
      cond_copy<N, MC>(dst, w.cond(res));
    }
  };
  int* const mine = req + 2 * l;
  if (role == 0) {
    int sf = 0, pr = 1, sp = 2;
    mine[0] = 0;
    mine[1] = -1;
    __syncthreads();
    int last[2] = {0, -1};
    for (int k = 1; k < K; ++k) {
      LmlStepVals<N, MC> v;
      if (active) lane_lml_step_part1<N, D, TPL>(c, q0, sf, pr, a.times[k] - a.times[k - 1], v, defer, stage_tail);
      __syncthreads();
      if (ckpt_every > 0 && active && (k - 1) % ckpt_every == 0) {
        // part 1 wrote only pr; the merge into sf's conditional is complete
        double* dst = ckpt + (size_t)((k - 1) / ckpt_every) * SlotLayout<N, MC>::size * n_lanes;
#pragma unroll
        for (int e = 0; e < SlotLayout<N, MC>::size; ++e) dst[(size_t)e * n_lanes] = w.at(sf, e);
      }
      last[0] = 0;
      last[1] = -1;
      if (active) {
        lane_lml_step_part2<N, D, TPL>(c, q0, sf, pr, v);
        const int ident = w.at(sf, LS_IDENT) != 0.0;
        last[0] = 1 | (ident << 1) | (sf << 2) | (pr << 4);
        last[1] = k % every == 0 ? k / every - 1 : -1;
        const int old_sf = sf;
        sf = pr;
        pr = sp;
        sp = old_sf;
        if (k % every == 0) w.at(sf, LS_IDENT) = 1.0;  // restart behind an identity
      }
      mine[0] = last[0];
      mine[1] = last[1];
      __syncthreads();
    }
    run(last);
    return sf;
  }
  __syncthreads();
  int sf = 0, pr = 1, sp = 2;
  for (int k = 1; k < K; ++k) {
    run(mine);
    __syncthreads();
    if (defer && active) {
      double Xm[N * MC];
#pragma unroll
      for (int i = 0; i < N * MC; ++i) Xm[i] = w.at(sf, SlotLayout<N, MC>::m + i);
      lane_revert_right<N, MC>(Xm, StageRef<N, MC>{w.cond(pr).p, stage_tail, kLaneBlock}, w.cond(pr), false);
    }
    if (active) {
      const int old_sf = sf;
      sf = pr;
      pr = sp;
      sp = old_sf;
    }
    __syncthreads();
  }
  return 0;
}
#endif

// Lane g of instance `inst`: the forward pass, the reverse scan, and its
// share of the outputs.  `slots` is the lane's LDS column (stride 64);
// `glane` / `n_lanes` index the lane in the HBM scratch buffer.  With SPLIT
// the forward pass runs on two wavefronts (lane_lml_forward_split); `active`
// tells whether the lane has work (every lane of both wavefronts calls this,
// the merging wavefront's lanes only for the forward pass), `lds_req` is the
// workgroup's request array.
template <int N, int D, int TPL, bool SPLIT = false>
__device__ __forceinline__ void lane_lml_instance(const LArgs& la, const LmlArgs& a, int inst, int g, double* slots,
                                                  size_t glane, size_t n_lanes, bool active = true,
                                                  int* lds_req = nullptr) {
  constexpr int MC = D * (1 + TPL);
  using CL = CondLayout<N, MC>;
  const pdq_config& cfg = a.cfg;
  LaneCtx<N, MC> c;
  c.a = &la;
  c.inst = inst;
  c.prm = a.params + (size_t)inst * cfg.n_params;
  c.w.base = slots;
  c.lin_set = 0;
  const LaneWs<N, MC>& w = c.w;
  const int ntan = a.n_tan, q0 = g * TPL;
  const int K = a.K, every = a.every, Kobs = (K - 1) / every + 1;
  int sf = 0, pr = 1;
  if (active && !(SPLIT && threadIdx.x >= kLaneBlock)) {
    // prior: mean = Taylor coefficients (with tangents), chol = init_std I
    const double* tc = a.tcoeffs + (size_t)inst * N * D * (1 + ntan);
    double m[N * MC], L[N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        m[i * MC + k] = tc[(i * D + k) * (1 + ntan)];
#pragma unroll
        for (int tt = 0; tt < TPL; ++tt) m[i * MC + (1 + tt) * D + k] = tc[(i * D + k) * (1 + ntan) + 1 + q0 + tt];
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) L[i * N + j] = i == j ? (cfg.init_std_rows ? cfg.init_std_rows[i] : cfg.init_std) : 0.0;
#pragma unroll
    for (int e = 0; e < SlotLayout<N, MC>::size; ++e) w.at(sf, e) = 0.0;
    w.store_mL(sf, m, L);
    w.at(sf, LS_T) = a.times[0];
    w.at(sf, LS_OS) = 1.0;
    w.at(sf, LS_IDENT) = 1.0;
  }
  if constexpr (SPLIT) {
#ifndef PDQ_HOST_EMU
    sf = lane_lml_forward_split<N, D, TPL>(c, a, q0, glane, n_lanes, active, lds_req);
    if (threadIdx.x >= kLaneBlock) return;  // the merging wavefront's part ends here
#endif
  } else {
    for (int k = 1; k < K; ++k) {
      lane_lml_step<N, D, TPL>(c, q0, sf, pr, a.times[k] - a.times[k - 1]);
      const int tmp = sf;
      sf = pr;
      pr = tmp;
      if (k % every == 0) {
        // the reset: store the interval's conditional, restart behind an identity
        const CondStrided<N, MC> dst{a.scratch + (size_t)(k / every - 1) * CL::size * n_lanes + glane, n_lanes};
        cond_copy<N, MC>(dst, w.cond(sf));
        w.at(sf, LS_IDENT) = 1.0;
      }
    }
  }
  if (!active) return;
  // terminal marginal: posterior1.conditional.marginalise(posterior1.marginal)
  double Xm[N * MC], XL[N * N];
  {
    double Sm[N * MC], SLf[N * N];
    w.load_mL(sf, Sm, SLf);
    if (w.at(sf, LS_IDENT) != 0.0) {
#pragma unroll
      for (int e = 0; e < N * MC; ++e) Xm[e] = Sm[e];
#pragma unroll
      for (int e = 0; e < N * N; ++e) XL[e] = SLf[e];
    } else {
      lane_marginalise_cond<N, MC>(w.cond(sf), Sm, SLf, Xm, XL);
    }
  }
  double lp[1 + TPL];
#pragma unroll
  for (int q = 0; q <= TPL; ++q) lp[q] = 0.0;
  double num = 0.0;
  const double c2pi = log(M_PI * 2.0);
  constexpr int n2 = N + 1;
  for (int j = Kobs - 1; j >= 0; --j) {
    if (j < Kobs - 1) {
      const CondStrided<N, MC> src{a.scratch + (size_t)j * CL::size * n_lanes + glane, n_lanes};
      double Ym[N * MC], YL[N * N];
      lane_marginalise_cond<N, MC>(src, Xm, XL, Ym, YL);
#pragma unroll
      for (int e = 0; e < N * MC; ++e) Xm[e] = Ym[e];
#pragma unroll
      for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
    }
    PDQ_PHASE_BOUNDARY();
    // model_j = marginal.to_derivative(0, std_j); bayes_rule_and_logpdf_tree
    const double* y = a.data + ((size_t)inst * Kobs + j) * D;
    const double sd = a.stds[(size_t)inst * Kobs + j];
    double big[n2 * n2], mo[MC];
#pragma unroll
    for (int r = 0; r < n2; ++r)
#pragma unroll
      for (int col = 0; col < n2; ++col) {
        double v;
        if (r < 1) v = col < 1 ? sd : 0.0;
        else v = col < 1 ? XL[0 * N + (r - 1)] : XL[(col - 1) * N + (r - 1)];
        big[r * n2 + col] = v;
      }
#pragma unroll
    for (int col = 0; col < MC; ++col) mo[col] = Xm[0 * MC + col] + 0.0;
    lane_qr<n2, n2, ObsRows0>(big);
    const double r00 = big[0], rinv = 1.0 / r00;
    double gain[N];
#pragma unroll
    for (int i = 0; i < N; ++i) gain[i] = lane_div(big[1 + i], r00, rinv);
    // observed.logpdf (isotropic: every data column against the same factor)
    double lpb[1 + TPL];
#pragma unroll
    for (int q = 0; q <= TPL; ++q) lpb[q] = 0.0;
    const double slogdet = log(fabs(r00));
#pragma unroll
    for (int col = 0; col < D; ++col) {
      const double wv = lane_div(y[col] - mo[col], r00, rinv);
      const double sq = 0.0 + wv * wv;
      lpb[0] += -0.5 * (2.0 * slogdet + sq + 1.0 * c2pi);
#pragma unroll
      for (int tt = 0; tt < TPL; ++tt) {
        const double wt = lane_div(-mo[(1 + tt) * D + col], r00, rinv);
        const double dot = 0.0 + wv * wt;
        lpb[1 + tt] += -dot;
      }
    }
#pragma unroll
    for (int q = 0; q <= TPL; ++q) {
      const double v = lpb[q];
      if (num == 0.0) lp[q] = v;
      else if (a.average_pdfs) lp[q] = (lp[q] * num + v) / (num + 1);
      else lp[q] = lp[q] + v;
    }
    num += 1.0;
    // updated = reverted.apply_flat(data)
    double Ym[N * MC], YL[N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < MC; ++col) {
        const double yo = col < D ? y[col] : 0.0;
        const double gy = 0.0 + gain[i] * yo;
        const double gm = 0.0 + gain[i] * mo[col];
        Ym[i * MC + col] = gy + (Xm[i * MC + col] - gm);
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k) YL[i * N + k] = big[(1 + k) * n2 + 1 + i];
#pragma unroll
    for (int e = 0; e < N * MC; ++e) Xm[e] = Ym[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
  }
  if (g == 0 && a.value) a.value[inst] = lp[0];
  if (a.grad)
#pragma unroll
    for (int tt = 0; tt < TPL; ++tt) a.grad[(size_t)inst * ntan + q0 + tt] = lp[1 + tt];
}

// LDS bytes of one workgroup: two slots per lane, or with the merging
// wavefront (SPLIT) three slots and 64 requests of two ints.
template <int N, int MC, bool SPLIT = false>
constexpr size_t lane_lml_lds_bytes() {
  return sizeof(double) * kLaneBlock * (SPLIT ? 3 : 2) * SlotLayout<N, MC>::size +
         (SPLIT ? sizeof(double) * kLaneBlock : 0);
}

// Whether the SPLIT form's LDS fits a workgroup (160 KB on gfx950).
template <int N, int MC>
constexpr bool lane_lml_split_fits() {
  return lane_lml_lds_bytes<N, MC, true>() <= 160 * 1024;
}

#ifndef PDQ_HOST_EMU
// One lane per (instance, tangent group), grid-stride over the n_inst G
// lanes, or with SPLIT one lane of wavefront 0 per (instance, tangent group)
// and the forward pass's merges on wavefront 1 (lane_lml_forward_split).
// The solver is a compile-time constant of the instantiation; the
// linearisation is ts0 (the C ABI refuses ts1 for the LML).
template <int N, int D, int TPL, int SOLVER, bool SPLIT>
__global__ void __launch_bounds__(SPLIT ? 2 * kLaneBlock : kLaneBlock) lane_lml_kernel(LmlArgs args) {
  extern __shared__ double lds[];
  LmlArgs a = args;
  a.cfg.strategy = PDQ_FIXEDPOINT;
  a.cfg.solver = SOLVER;
  a.cfg.constraint = PDQ_TS0;
  LArgs la{};
  la.cfg = a.cfg;
  la.n_inst = a.n_inst;
  la.K = a.K;
  la.Q = a.Q;
  const size_t n_lanes = (size_t)a.n_inst * a.G;
  if constexpr (SPLIT) {
    constexpr int MC = D * (1 + TPL);
    const int l = threadIdx.x & (kLaneBlock - 1);
    const size_t gl = (size_t)blockIdx.x * kLaneBlock + l;
    const bool active = gl < n_lanes;
    const size_t g2 = active ? gl : 0;
    int* const req = reinterpret_cast<int*>(lds + kLaneBlock * 3 * SlotLayout<N, MC>::size);
    lane_lml_instance<N, D, TPL, true>(la, a, (int)(g2 / a.G), (int)(g2 % a.G), lds + l, gl, n_lanes, active, req);
  } else {
    double* const slots = lds + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t gl = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gl < n_lanes; gl += stride)
      lane_lml_instance<N, D, TPL>(la, a, (int)(gl / a.G), (int)(gl % a.G), slots, gl, n_lanes);
  }
}
#endif

}  // namespace pdq
