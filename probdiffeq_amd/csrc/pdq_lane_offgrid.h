// Off-grid marginals of a solution (ProbabilisticSolver.offgrid_marginals,
// solvers.py:149-203), isotropic, one lane per query time.
//
// For a query t the lane finds the grid interval (t0, t1] by
// searchsorted(t_grid, t) (side "left", as the reference) and evaluates
// * the filter (estimators_and_losses.py:403-414): the marginal at t0
//   predicted to t with the transition of t1's output scale;
// * the fixed-interval smoother (:674-707): the filtering marginal at t0
//   reverted to t and on to t1; the smoothed marginal at t1 marginalised
//   back through the conditional t1 -> t.
// The per-query algebra is the lane solve's (pdq_lane.h), all in registers.
// Like the reference, t must lie strictly inside the grid and off its
// points; other t give an unspecified result (the interval index is clamped
// to [1, K - 1], so every access stays inside the solution arrays).
#pragma once

#include "pdq_lane.h"

namespace pdq {

template <int N, int D, int STRATEGY>
__device__ __forceinline__ void lane_offgrid(const OffgridArgs& a, size_t q) {
  const pdq_config& cfg = a.cfg;
  const int K = a.K;
  const size_t inst = q / (size_t)a.n_query;
  const double* __restrict__ ts = a.times + inst * K;
  const double t = a.t_query[q];
  int lo = 0, hi = K;
  while (lo < hi) {  // first index with ts[i] >= t: at most log2(K) + 1 rounds
    const int mid = (lo + hi) / 2;
    if (ts[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  const int i = lo < 1 ? 1 : (lo > K - 1 ? K - 1 : lo);
  const double t0 = ts[i - 1], t1 = ts[i];
  const double os = a.output_scale[inst * K + i];
  const size_t r0 = inst * K + (i - 1), r1 = inst * K + i;
  const double* m0p = (STRATEGY == PDQ_FILTER ? a.mean : a.filt_mean) + r0 * N * D;
  const double* L0p = (STRATEGY == PDQ_FILTER ? a.chol : a.filt_chol) + r0 * N * N;
  double m0[N * D], L0[N * N], P[N], PI[N], Ym[N * D], YL[N * N];
#pragma unroll
  for (int e = 0; e < N * D; ++e) m0[e] = m0p[e];
#pragma unroll
  for (int e = 0; e < N * N; ++e) L0[e] = L0p[e];
  lane_precon<N>(t - t0, P, PI);
  const double s0 = lane_noise_scale(cfg, t - t0, os);
  if constexpr (STRATEGY == PDQ_FILTER) {
    lane_marginalise_prior<N, D>(a.Q, P, PI, s0, m0, L0, Ym, YL);
  } else {
    double Xm[N * D], XL[N * N], Zm[N * D], ZL[N * N], m1[N * D], L1[N * N];
    LCond<N, D> c_t_t0, c_t1_t;
    lane_revert_prior<N, D>(a.Q, P, PI, s0, m0, L0, Xm, XL, c_t_t0);
    lane_precon<N>(t1 - t, P, PI);
    const double s1 = lane_noise_scale(cfg, t1 - t, os);
    lane_revert_prior<N, D>(a.Q, P, PI, s1, Xm, XL, Zm, ZL, c_t1_t);
#pragma unroll
    for (int e = 0; e < N * D; ++e) m1[e] = a.mean[r1 * N * D + e];
#pragma unroll
    for (int e = 0; e < N * N; ++e) L1[e] = a.chol[r1 * N * N + e];
    lane_marginalise_cond<N, D>(c_t1_t, m1, L1, Ym, YL);
  }
#pragma unroll
  for (int e = 0; e < N * D; ++e) a.mean_out[q * N * D + e] = Ym[e];
#pragma unroll
  for (int e = 0; e < N * N; ++e) a.chol_out[q * N * N + e] = YL[e];
}

#ifndef PDQ_HOST_EMU
template <int N, int D, int STRATEGY>
__global__ void __launch_bounds__(64) lane_offgrid_kernel(OffgridArgs a) {
  const size_t total = (size_t)a.n_inst * a.n_query;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride)
    lane_offgrid<N, D, STRATEGY>(a, q);
}
#endif

}  // namespace pdq
