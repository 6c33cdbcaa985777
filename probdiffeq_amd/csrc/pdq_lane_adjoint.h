// Reverse-mode adjoint of the lane-per-instance marginal log-likelihood
// (isotropic, ts0): the gradient of loss_lml_timeseries (estimators_and_
// losses.py:53-105) on the fixed-point smoother's Markov sequence with
// respect to the vector field's parameters and the initial value, by one
// lane per instance.
//
// What carries derivatives.  With ts0 and the integrated-Wiener-process prior
// every square-root factor of the forward pass and of the reverse scan is a
// function of the grid, the prior, the observation stds and (for
// solver_dynamic) the output scales, which the reference differentiates as
// constants (stop_gradient, solvers.py:568-575).  So the adjoint of every
// R-only QR (backend/linalg.py:8-37) with respect to the parameters is zero,
// and the reverse pass propagates adjoints through the mean recursions alone:
// the transition (revert), the fixed-point merge (estimators_and_losses.py:
// 473-591), the update (ssm_impl_api.py:112-116) with its linearisation
// b = -f(m_0) (whose vector-field Jacobians bring in the parameters), and
// MarkovSequence.evaluate_lml's reverse scan (:180-218).
//
// Schedule (checkpointing and recomputation).  The forward pass stores the
// solution slot every C grid steps (C = 50) and the backward conditional of
// every observation interval; the reverse scan stores each observation
// node's gain, observed factor and whitened residuals.  The scan's adjoint
// then runs forward over the nodes and yields the adjoint of every interval
// conditional's noise mean and of the terminal marginal.  The backward sweep
// visits the checkpoint segments from last to first: it recomputes the
// segment's slots from its checkpoint, then walks the segment backwards,
// recomputing each step's factors and gains from the stored slot (the
// forward factors are recomputed, not stored) and applying the step's
// adjoint.  The adjoint of the initial mean meets the Taylor coefficients'
// sensitivities (forward-mode duals of the jet recursion, 6 directions).
//
// Storage per lane, in HBM element-major across the launch's lanes:
// conditionals (Kobs - 1) x |cond|, nodes Kobs x (N + 1 + D), adjoints of the
// conditional means (Kobs - 1) x N D, checkpoints ceil((K - 1) / C) x |slot|,
// and two segment buffers of C slots.
//
// Two-wavefront form (SPLIT).  The forward pass is the LML kernel's split one
// (lane_lml_forward_split: wavefront 1 runs each step's merge and the copy of
// each observation interval's conditional beside wavefront 0's next step;
// three slots rotate, the third in wavefront 1's region of the LDS, and
// wavefront 0 stores the checkpoints).  Then wavefront 1, on its own two LDS
// slots, recomputes segment sg - 1 into one segment buffer while wavefront 0
// walks segment sg backwards from the other, and recomputes the last segment
// while wavefront 0 runs the reverse scan and its adjoint.  The workgroup
// barriers sit at phase boundaries that every lane reaches: lanes without an
// instance of their own compute an existing one and write no output.
#pragma once

#include "pdq_lane_lml.h"

namespace pdq {

constexpr int kAdjSegment = 50;  // grid steps per checkpoint segment

// Offsets (in doubles per lane) of the adjoint's storage regions.
template <int N, int D>
struct AdjLayout {
  using CL = CondLayout<N, D>;
  using SL = SlotLayout<N, D>;
  static constexpr int node = N + 1 + D;  // gain, r00, whitened residuals
  static constexpr bool conds_first = true;  // conds sits at offset 0 (the split forward pass relies on it)
  size_t conds, nodes, xibar, ckpt, seg, size;
  __device__ __host__ AdjLayout(int K, int every) {
    const size_t kobs = (size_t)(K - 1) / every + 1;
    const size_t nck = (size_t)(K - 2) / kAdjSegment + 1;
    size_t o = 0;
    conds = o;
    o += (kobs > 1 ? kobs - 1 : 1) * CL::size;
    nodes = o;
    o += kobs * node;
    xibar = o;
    o += (kobs > 1 ? kobs - 1 : 1) * N * D;
    ckpt = o;
    o += nck * SL::size;
    seg = o;
    o += 2 * (size_t)kAdjSegment * SL::size;
    size = o;
  }
};

// A register conditional (N x N, N x D mean) with CondRef's writable
// accessors, for lane_revert_prior.
template <int N, int D>
struct LCondViewD {
  using CL = CondLayout<N, D>;
  LCond<N, D>* p;
  __device__ __forceinline__ double& at(int e) const { return p->v[e]; }
  __device__ __forceinline__ double& A(int i, int j) const { return p->v[CL::A + i * N + j]; }
  __device__ __forceinline__ double& m(int i, int c) const { return p->v[CL::m + i * D + c]; }
  __device__ __forceinline__ double& tl(int i) const { return p->v[CL::tl + i]; }
  __device__ __forceinline__ double& to(int i) const { return p->v[CL::to + i]; }
  __device__ __forceinline__ double L(int i, int j) const { return j <= i ? p->v[CL::L + tri(i, j)] : 0.0; }
  __device__ __forceinline__ void setL(int i, int j, double x) const {
    if (j <= i) p->v[CL::L + tri(i, j)] = x;
  }
};

// Copies a lane's LDS slot to / from HBM storage (element e at p[e * stride]).
template <int N, int D>
__device__ __forceinline__ void adj_slot_out(const LaneWs<N, D>& w, int s, double* p, size_t stride) {
#pragma unroll
  for (int e = 0; e < SlotLayout<N, D>::size; ++e) p[(size_t)e * stride] = w.at(s, e);
}
template <int N, int D>
__device__ __forceinline__ void adj_slot_in(const LaneWs<N, D>& w, int s, const double* p, size_t stride) {
#pragma unroll
  for (int e = 0; e < SlotLayout<N, D>::size; ++e) w.at(s, e) = p[(size_t)e * stride];
}

// Recomputes checkpoint segment sg's slots (grid steps k0 .. k1 - 1, from the
// checkpoint stored by the forward pass) into segment buffer `buf`, stepping
// on the lane's LDS slots 0 and 1 as the forward pass did.
template <int N, int D, class At>
__device__ __forceinline__ void adj_recompute_segment(const LaneCtx<N, D>& c, const LmlArgs& a,
                                                      const AdjLayout<N, D>& lay, At at, size_t n_lanes, int sg,
                                                      int buf) {
  using SL = SlotLayout<N, D>;
  const LaneWs<N, D>& w = c.w;
  const int K = a.K, every = a.every;
  const int k0 = sg * kAdjSegment;
  const int k1 = (k0 + kAdjSegment < K - 1) ? k0 + kAdjSegment : K - 1;
  int sf = 0, pr = 1;
  adj_slot_in<N, D>(w, sf, at(lay.ckpt + (size_t)sg * SL::size), n_lanes);
  for (int k = k0; k < k1; ++k) {
    adj_slot_out<N, D>(w, sf, at(lay.seg + ((size_t)buf * kAdjSegment + (k - k0)) * SL::size), n_lanes);
    if (k + 1 == k1) break;
    lane_lml_step<N, D, 0>(c, 0, sf, pr, a.times[k + 1] - a.times[k]);
    const int tmp = sf;
    sf = pr;
    pr = tmp;
    if ((k + 1) % every == 0) w.at(sf, LS_IDENT) = 1.0;
  }
}

// The quantities of one forward step (from slot `from`, step h) that its
// adjoint needs, recomputed from the slot: the transition's preconditioner,
// the revert's gain G_b and its output scales to_b, the prediction mean Pm
// and the update's gain (solver.step / solver_dynamic.step, solvers.py:
// 549-599, 702-733, with the operations of lane_lml_step).
template <int N, int D>
struct AdjStep {
  double P[N], PI[N], Gb[N * N], tob[N], Pm[N * D], gain[N];
};

template <int N, int D>
__device__ __forceinline__ void adj_recompute_step(const LaneCtx<N, D>& c, int from, double h, AdjStep<N, D>& s) {
  const pdq_config& cfg = c.a->cfg;
  const double* __restrict__ Q = c.a->Q;
  const LaneWs<N, D>& w = c.w;
  double Sm[N * D], SLf[N * N];
  w.load_mL(from, Sm, SLf);
  const double tn = w.at(from, LS_T) + h;
  lane_precon<N>(h, s.P, s.PI);
  double sigma = 1.0, b[D];
  if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
    const double s1 = lane_noise_scale(cfg, h, 1.0);
    double Um[N * D], col[N + 1];
    lane_apply_flat_mean<N, D>(s.P, s.PI, Sm, Um);
    lane_lml_linearize<N, D, 0>(cfg, c.prm, 0, Um, tn, b);
#pragma unroll
    for (int r = 0; r < N; ++r) col[r] = fabs(s.P[1]) * (s1 * lane_iwp_q<N>(Q, 1, r));
    col[N] = cfg.damp;
    double om[D];
#pragma unroll
    for (int k = 0; k < D; ++k) om[k] = Um[1 * D + k] + b[k];
    lane_qr<N + 1, 1, FxColumnRows<N>>(col);
    sigma = lane_rms_zero<D>(om, col[0]);
  }
  LCond<N, D> bw;
  double PL[N * N];
  lane_revert_prior<N, D>(Q, s.P, s.PI, lane_noise_scale(cfg, h, sigma), Sm, SLf, s.Pm, PL, LCondViewD<N, D>{&bw});
#pragma unroll
  for (int e = 0; e < N * N; ++e) s.Gb[e] = bw.v[CondLayout<N, D>::A + e];
#pragma unroll
  for (int i = 0; i < N; ++i) s.tob[i] = bw.to(i);
  // the update's gain: [[damp, 0], [(H PL)^T, PL^T]] with H = e_1 (lane_update)
  constexpr int n2 = N + 1;
  double big[n2 * n2];
#pragma unroll
  for (int r = 0; r < n2; ++r)
#pragma unroll
    for (int col = 0; col < n2; ++col) {
      double v;
      if (r < 1) v = col < 1 ? cfg.damp : 0.0;
      else v = col < 1 ? PL[1 * N + (r - 1)] : PL[(col - 1) * N + (r - 1)];
      big[r * n2 + col] = v;
    }
  lane_qr<n2, n2, UpdateRows>(big);
  const double r0 = 1.0 / big[0];
#pragma unroll
  for (int j = 0; j < N; ++j) s.gain[j] = lane_div(big[1 + j], big[0], r0);
}

// One lane: the forward pass with checkpoints, the reverse scan (value), the
// scan's adjoint, and the backward sweep over the checkpoint segments.  With
// SPLIT this is wavefront 0's part (lane_adj_recompute_role is wavefront 1's),
// `req` the workgroup's request array of the split forward pass; `active` is
// false for a lane that computes an instance it does not own (it writes no
// output).
template <int N, int D, bool SPLIT = false>
__device__ __forceinline__ void lane_adj_instance(const LArgs& la, const LmlArgs& a, int inst, double* slots,
                                                  size_t lane, size_t n_lanes, bool active = true,
                                                  int* req = nullptr, double* stage_tail = nullptr) {
  using CL = CondLayout<N, D>;
  using SL = SlotLayout<N, D>;
  const pdq_config& cfg = a.cfg;
  LaneCtx<N, D> c;
  c.a = &la;
  c.inst = inst;
  c.prm = a.params + (size_t)inst * cfg.n_params;
  c.w.base = slots;
  c.lin_set = 0;
  const LaneWs<N, D>& w = c.w;
  const int K = a.K, every = a.every, Kobs = (K - 1) / every + 1, ntan = a.n_tan;
  const AdjLayout<N, D> lay(K, every);
  double* const base = a.scratch + lane;
  auto at = [&](size_t off) -> double* { return base + off * n_lanes; };
  int sf = 0, pr = 1;
  {
    // prior: mean = Taylor coefficients, chol = init_std I
    const double* tc = a.tcoeffs + (size_t)inst * N * D * (1 + ntan);
    double m[N * D], L[N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) m[i * D + k] = tc[(i * D + k) * (1 + ntan)];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) L[i * N + j] = i == j ? (cfg.init_std_rows ? cfg.init_std_rows[i] : cfg.init_std) : 0.0;
#pragma unroll
    for (int e = 0; e < SL::size; ++e) w.at(sf, e) = 0.0;
    w.store_mL(sf, m, L);
    w.at(sf, LS_T) = a.times[0];
    w.at(sf, LS_OS) = 1.0;
    w.at(sf, LS_IDENT) = 1.0;
  }
  // ---- forward pass: checkpoints every kAdjSegment steps, interval conditionals
  if constexpr (SPLIT) {
#ifndef PDQ_HOST_EMU
    // the interval conditionals land where the LML kernel puts them: lay.conds
    // is the storage's first region
    static_assert(AdjLayout<N, D>::conds_first, "the split forward pass writes the conditionals at offset 0");
    sf = lane_lml_forward_split<N, D, 0>(c, a, 0, lane, n_lanes, true, req, kAdjSegment, at(lay.ckpt), stage_tail);
#endif
  } else {
    for (int k = 0; k < K - 1; ++k) {
      if (k % kAdjSegment == 0) adj_slot_out<N, D>(w, sf, at(lay.ckpt + (size_t)(k / kAdjSegment) * SL::size), n_lanes);
      lane_lml_step<N, D, 0>(c, 0, sf, pr, a.times[k + 1] - a.times[k]);
      const int tmp = sf;
      sf = pr;
      pr = tmp;
      if ((k + 1) % every == 0) {
        const CondStrided<N, D> dst{at(lay.conds + (size_t)((k + 1) / every - 1) * CL::size), n_lanes};
        cond_copy<N, D>(dst, w.cond(sf));
        w.at(sf, LS_IDENT) = 1.0;
      }
    }
  }
  // terminal marginal: posterior1.conditional.marginalise(posterior1.marginal)
  const bool term_ident = w.at(sf, LS_IDENT) != 0.0;
  double Xm[N * D], XL[N * N];
  double tA[N * N], ttl[N], tto[N];  // the final running conditional (terminal adjoint)
  {
    double Sm[N * D], SLf[N * N];
    w.load_mL(sf, Sm, SLf);
    const CondRef<N, D> fc = w.cond(sf);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      ttl[i] = fc.tl(i);
      tto[i] = fc.to(i);
#pragma unroll
      for (int j = 0; j < N; ++j) tA[i * N + j] = fc.A(i, j);
    }
    if (term_ident) {
#pragma unroll
      for (int e = 0; e < N * D; ++e) Xm[e] = Sm[e];
#pragma unroll
      for (int e = 0; e < N * N; ++e) XL[e] = SLf[e];
    } else {
      lane_marginalise_cond<N, D>(fc, Sm, SLf, Xm, XL);
    }
  }
  if constexpr (SPLIT) {
    // every checkpoint is stored and the terminal slot read (with SPLIT it may
    // be the third slot, in wavefront 1's region): wavefront 1 starts
    // recomputing; wavefront 0 walks on slots 0 and 1 from here on
    __syncthreads();
    sf = 0;
  }
  // ---- reverse scan (evaluate_lml): value, and each node's gain, factor, residuals
  double lp = 0.0, num = 0.0;
  const double c2pi = log(M_PI * 2.0);
  constexpr int n2 = N + 1;
  for (int j = Kobs - 1; j >= 0; --j) {
    if (j < Kobs - 1) {
      const CondStrided<N, D> src{at(lay.conds + (size_t)j * CL::size), n_lanes};
      double Ym[N * D], YL[N * N];
      lane_marginalise_cond<N, D>(src, Xm, XL, Ym, YL);
#pragma unroll
      for (int e = 0; e < N * D; ++e) Xm[e] = Ym[e];
#pragma unroll
      for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
    }
    PDQ_PHASE_BOUNDARY();
    const double* y = a.data + ((size_t)inst * Kobs + j) * D;
    const double sd = a.stds[(size_t)inst * Kobs + j];
    double big[n2 * n2], mo[D];
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
    for (int col = 0; col < D; ++col) mo[col] = Xm[0 * D + col] + 0.0;
    lane_qr<n2, n2, ObsRows0>(big);
    const double r00 = big[0], rinv = 1.0 / r00;
    double gain[N];
#pragma unroll
    for (int i = 0; i < N; ++i) gain[i] = lane_div(big[1 + i], r00, rinv);
    double lpb = 0.0, wv[D];
    const double slogdet = log(fabs(r00));
#pragma unroll
    for (int col = 0; col < D; ++col) {
      wv[col] = lane_div(y[col] - mo[col], r00, rinv);
      const double sq = 0.0 + wv[col] * wv[col];
      lpb += -0.5 * (2.0 * slogdet + sq + 1.0 * c2pi);
    }
    if (num == 0.0) lp = lpb;
    else if (a.average_pdfs) lp = (lp * num + lpb) / (num + 1);
    else lp = lp + lpb;
    num += 1.0;
    double* nd = at(lay.nodes + (size_t)j * AdjLayout<N, D>::node);
#pragma unroll
    for (int i = 0; i < N; ++i) nd[(size_t)i * n_lanes] = gain[i];
    nd[(size_t)N * n_lanes] = r00;
#pragma unroll
    for (int col = 0; col < D; ++col) nd[(size_t)(N + 1 + col) * n_lanes] = wv[col];
    // updated = reverted.apply_flat(data)
    double Ym[N * D], YL[N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < D; ++col) {
        const double gy = 0.0 + gain[i] * y[col];
        const double gm = 0.0 + gain[i] * mo[col];
        Ym[i * D + col] = gy + (Xm[i * D + col] - gm);
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k) YL[i * N + k] = big[(1 + k) * n2 + 1 + i];
#pragma unroll
    for (int e = 0; e < N * D; ++e) Xm[e] = Ym[e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) XL[e] = YL[e];
  }
  if (a.value && active) a.value[inst] = lp;
  if (!a.grad) return;
  // ---- adjoint of the scan, forward over the nodes
  const double cw = a.average_pdfs ? 1.0 / (double)Kobs : 1.0;  // d lml / d logpdf_j
  double corr_bar[N * D], xterm_bar[N * D];
#pragma unroll
  for (int e = 0; e < N * D; ++e) corr_bar[e] = 0.0;
  for (int j = 0; j < Kobs; ++j) {
    const double* nd = at(lay.nodes + (size_t)j * AdjLayout<N, D>::node);
    double gain[N], wv[D];
#pragma unroll
    for (int i = 0; i < N; ++i) gain[i] = nd[(size_t)i * n_lanes];
    const double r00 = nd[(size_t)N * n_lanes];
#pragma unroll
    for (int col = 0; col < D; ++col) wv[col] = nd[(size_t)(N + 1 + col) * n_lanes];
    // corrected_j = pred_j + gain (y - pred_j[0]); logpdf_j = -1/2 |(y - pred_j[0]) / r00|^2 + ...
    double pred_bar[N * D];
#pragma unroll
    for (int col = 0; col < D; ++col) {
      double mo_bar = cw * wv[col] / r00;
#pragma unroll
      for (int i = 0; i < N; ++i) mo_bar -= gain[i] * corr_bar[i * D + col];
#pragma unroll
      for (int i = 0; i < N; ++i) pred_bar[i * D + col] = corr_bar[i * D + col] + (i == 0 ? mo_bar : 0.0);
    }
    if (j == Kobs - 1) {
#pragma unroll
      for (int e = 0; e < N * D; ++e) xterm_bar[e] = pred_bar[e];
      break;
    }
    // pred_j = to o (A (tl o corrected_{j+1}) + xi_j)  (marginalise of conditional j)
    const CondStrided<N, D> cj{at(lay.conds + (size_t)j * CL::size), n_lanes};
    double* xb = at(lay.xibar + (size_t)j * N * D);
    double tp[N * D];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < D; ++col) {
        tp[i * D + col] = cj.to(i) * pred_bar[i * D + col];
        xb[(size_t)(i * D + col) * n_lanes] = tp[i * D + col];
      }
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int col = 0; col < D; ++col) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += cj.A(i, k) * tp[i * D + col];
        corr_bar[k * D + col] = cj.tl(k) * acc;
      }
  }
  // ---- backward sweep over the forward steps
  double mbar[N * D], cbar[N * D], tbar[kMaxParams > 8 ? 8 : kMaxParams];
  const int np = cfg.n_params < 8 ? cfg.n_params : 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) tbar[q] = 0.0;
  if (term_ident) {
#pragma unroll
    for (int e = 0; e < N * D; ++e) {
      mbar[e] = xterm_bar[e];
      cbar[e] = 0.0;
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int col = 0; col < D; ++col) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += tA[i * N + k] * (tto[i] * xterm_bar[i * D + col]);
        mbar[k * D + col] = ttl[k] * acc;
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int col = 0; col < D; ++col) cbar[i * D + col] = tto[i] * xterm_bar[i * D + col];
  }
  const int nseg = (K - 2) / kAdjSegment + 1;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int k0 = sg * kAdjSegment;
    const int k1 = (k0 + kAdjSegment < K - 1) ? k0 + kAdjSegment : K - 1;
    // the segment's slots from its checkpoint: recomputed here, or by
    // wavefront 1 into buffer sg & 1 before this barrier
    const int buf = SPLIT ? (sg & 1) : 0;
    if constexpr (SPLIT) __syncthreads();
    else adj_recompute_segment<N, D>(c, a, lay, at, n_lanes, sg, 0);
    for (int k = k1 - 1; k >= k0; --k) {
      // after this step a reset stored the running conditional as interval j's:
      // its adjoint comes from the scan, and the identity behind it has none
      if ((k + 1) % every == 0) {
        const double* xb = at(lay.xibar + (size_t)((k + 1) / every - 1) * N * D);
#pragma unroll
        for (int e = 0; e < N * D; ++e) cbar[e] = xb[(size_t)e * n_lanes];
      }
      adj_slot_in<N, D>(w, sf, at(lay.seg + ((size_t)buf * kAdjSegment + (k - k0)) * SL::size), n_lanes);
      AdjStep<N, D> s;
      adj_recompute_step<N, D>(c, sf, a.times[k + 1] - a.times[k], s);
      const bool ident = w.at(sf, LS_IDENT) != 0.0;
      // update: m_{k+1} = Pm - gain (Pm[1] + b), b = -f(Pm[0])
      double Pbar[N * D], bbar[D];
#pragma unroll
      for (int col = 0; col < D; ++col) {
        double mo_bar = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) mo_bar -= s.gain[i] * mbar[i * D + col];
        bbar[col] = mo_bar;
#pragma unroll
        for (int i = 0; i < N; ++i) Pbar[i * D + col] = mbar[i * D + col] + (i == 1 ? mo_bar : 0.0);
      }
      // linearisation: d(-f)/dy and d(-f)/dtheta at y = Pm[0]
      {
        double yv[D];
#pragma unroll
        for (int jj = 0; jj < D; ++jj) yv[jj] = s.Pm[jj];
#pragma unroll
        for (int jj = 0; jj < D; ++jj) {
          double acc = 0.0;
#pragma unroll
          for (int jf = 0; jf < D; ++jf) acc += bbar[jf] * vf_jac(cfg.vector_field, D, yv, c.prm, jf, jj);
          Pbar[jj] -= acc;
        }
        for (int q = 0; q < np; ++q) {
          double acc = 0.0;
#pragma unroll
          for (int jf = 0; jf < D; ++jf) acc += bbar[jf] * vf_jac_param(cfg.vector_field, D, yv, c.prm, jf, q);
          tbar[q] -= acc;
        }
      }
      // merge(self = running conditional, other = bw): xi' = A_self (T o xi_bw) + xi_self,
      // or to_bw o xi_bw behind an identity
      double mbbar[N * D];
      if (ident) {
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
          for (int col = 0; col < D; ++col) mbbar[i * D + col] = s.tob[i] * cbar[i * D + col];
#pragma unroll
        for (int e = 0; e < N * D; ++e) cbar[e] = 0.0;
      } else {
        const CondRef<N, D> self = w.cond(sf);
#pragma unroll
        for (int k2 = 0; k2 < N; ++k2) {
          const double T = self.tl(k2) * s.tob[k2];
#pragma unroll
          for (int col = 0; col < D; ++col) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) acc += self.A(i, k2) * cbar[i * D + col];
            mbbar[k2 * D + col] = T * acc;
          }
        }
      }
      // revert: tmp = A (PI o m); Pm = P o tmp; xi_bw = PI o m - G_b tmp
      double tbarv[N * D];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int col = 0; col < D; ++col) {
          double acc = s.P[i] * Pbar[i * D + col];
#pragma unroll
          for (int r = 0; r < N; ++r) acc -= s.Gb[r * N + i] * mbbar[r * D + col];
          tbarv[i * D + col] = acc;
        }
#pragma unroll
      for (int k2 = 0; k2 < N; ++k2)
#pragma unroll
        for (int col = 0; col < D; ++col) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const double av = lane_iwp_a<N>(i, k2);
            if (av != 0.0) acc += av * tbarv[i * D + col];
          }
          mbar[k2 * D + col] = s.PI[k2] * (mbbar[k2 * D + col] + acc);
        }
    }
  }
  // the initial mean's adjoint against the Taylor coefficients' sensitivities
  const double* tc = a.tcoeffs + (size_t)inst * N * D * (1 + ntan);
  for (int q = 0; q < ntan; ++q) {
    double g = q < np ? tbar[q] : 0.0;
#pragma unroll
    for (int e = 0; e < N * D; ++e) g += mbar[e] * tc[(size_t)e * (1 + ntan) + 1 + q];
    if (active) a.grad[(size_t)inst * ntan + q] = g;
  }
}

// Wavefront 1 of the two-wavefront adjoint: the merges of the split forward
// pass on wavefront 0's slots (`fwd_slots`), then the recomputation of the
// checkpoint segments, one segment ahead of wavefront 0's backward walk, on
// the lane's own LDS slots (`slots`), for the same instance and HBM column as
// its partner lane of wavefront 0.  Its barriers pair with
// lane_adj_instance<N, D, true>'s.
template <int N, int D>
__device__ __forceinline__ void lane_adj_recompute_role(const LArgs& la, const LmlArgs& a, int inst, double* fwd_slots,
                                                        double* slots, size_t lane, size_t n_lanes, int* req,
                                                        double* stage_tail) {
  LaneCtx<N, D> c;
  c.a = &la;
  c.inst = inst;
  c.prm = a.params + (size_t)inst * a.cfg.n_params;
  c.w.base = fwd_slots;
  c.lin_set = 0;
#ifndef PDQ_HOST_EMU
  lane_lml_forward_split<N, D, 0>(c, a, 0, lane, n_lanes, true, req, 0, nullptr, stage_tail);
#endif
  c.w.base = slots;
  const AdjLayout<N, D> lay(a.K, a.every);
  double* const base = a.scratch + lane;
  auto at = [&](size_t off) -> double* { return base + off * n_lanes; };
  const int nseg = (a.K - 2) / kAdjSegment + 1;
  __syncthreads();  // the checkpoints are stored and wavefront 0 has read the terminal slot
  if (!a.grad) return;
  adj_recompute_segment<N, D>(c, a, lay, at, n_lanes, nseg - 1, (nseg - 1) & 1);
  for (int sg = nseg - 1; sg >= 0; --sg) {
    __syncthreads();  // segment sg is in its buffer: wavefront 0 walks it now
    if (sg > 0) adj_recompute_segment<N, D>(c, a, lay, at, n_lanes, sg - 1, (sg - 1) & 1);
  }
}

#ifndef PDQ_HOST_EMU
// Bytes of LDS of the adjoint kernel: two slots per lane of each wavefront,
// and with SPLIT the forward pass's request array (two ints per lane).
template <int N, int D, bool SPLIT>
constexpr size_t lane_adj_lds_bytes() {
  return sizeof(double) * kLaneBlock * 2 * SlotLayout<N, D>::size * (SPLIT ? 2 : 1) +
         (SPLIT ? 2 * sizeof(int) * kLaneBlock + sizeof(double) * kLaneBlock * RevertStage<N, D>::tail : 0);
}

// One lane per instance (the grid covers every instance once); slots in LDS
// as in the forward LML kernel.  With SPLIT the workgroup is two wavefronts:
// lane l of wavefront 0 runs the adjoint of instance blockIdx.x * 64 + l and
// lane l of wavefront 1 recomputes its checkpoint segments on its own slots.
template <int N, int D, int SOLVER, bool SPLIT>
__global__ void __launch_bounds__(SPLIT ? 2 * kLaneBlock : kLaneBlock) lane_adj_kernel(LmlArgs args) {
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
  if constexpr (SPLIT) {
    const int l = threadIdx.x & (kLaneBlock - 1);
    const int role = __builtin_amdgcn_readfirstlane(threadIdx.x / kLaneBlock);
    // wavefront 0's slots 0, 1 (and, in the forward pass, 2), then wavefront
    // 1's slots 0, 1 (the forward pass's slot 2 among them), then the requests
    double* const fwd_slots = lds + l;
    double* const slots = lds + l + (role ? 2 * SlotLayout<N, D>::size * kLaneBlock : 0);
    int* const req = reinterpret_cast<int*>(lds + 4 * SlotLayout<N, D>::size * kLaneBlock);
    // after the requests: the lane's revert stage tail (the forward pass's
    // conditionals are finished on wavefront 1)
    double* const stage_tail =
        RevertStage<N, D>::tail > 0 ? reinterpret_cast<double*>(req + 2 * kLaneBlock) + l : nullptr;
    const size_t n_lanes = (size_t)gridDim.x * kLaneBlock;
    const size_t lane = (size_t)blockIdx.x * kLaneBlock + l;
    // a lane past the last instance computes the last one (uniform control
    // flow for the barriers) and writes nothing
    const bool active = lane < (size_t)a.n_inst;
    const int inst = active ? (int)lane : a.n_inst - 1;
    if (role == 0) lane_adj_instance<N, D, true>(la, a, inst, slots, lane, n_lanes, active, req, stage_tail);
    else lane_adj_recompute_role<N, D>(la, a, inst, fwd_slots, slots, lane, n_lanes, req, stage_tail);
  } else {
    double* const slots = lds + threadIdx.x;
    const size_t n_lanes = (size_t)gridDim.x * blockDim.x;
    const size_t lane = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t inst = lane; inst < (size_t)a.n_inst; inst += n_lanes) lane_adj_instance<N, D>(la, a, (int)inst, slots, lane, n_lanes);
  }
}
#endif

}  // namespace pdq
