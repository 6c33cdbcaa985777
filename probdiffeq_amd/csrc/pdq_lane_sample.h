// Samples of a smoother's posterior (MarkovSequence.sample,
// estimators_and_losses.py:233-273), isotropic, one lane per (instance,
// sample).
//
// The reverse scan starts from a draw of the terminal marginal and pushes
// each sample back through the backward conditionals: predicted =
// cond.apply_flat(x), x = predicted.sample_flat(z).  The isotropic forms are
// kept exactly (ssm_impl_isotropic.py:75-79, 255-258): apply_flat adds the
// noise mean without the to_observed scaling the dense form applies, and
// sample_flat draws one N-vector z that every one of the d columns shares.
// The standard normal draws come in as data (the reference draws them from
// JAX keys): base[instance, sample, k] for conditional k, row K - 1 for the
// terminal marginal.
#pragma once

#include "pdq_lane.h"

namespace pdq {

template <int N, int D>
__device__ __forceinline__ void lane_sample(const SampleArgs& a, size_t gid) {
  const int K = a.K;
  const size_t inst = gid / (size_t)a.n_samples;
  const double* __restrict__ z = a.base + gid * K * N;
  double* __restrict__ out = a.samples + gid * K * N * D;
  double x[N * D];
  {
    const double* m = a.term_mean + inst * N * D;
    const double* L = a.term_chol + inst * N * N;
    const double* zk = z + (size_t)(K - 1) * N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) acc += L[i * N + j] * zk[j];
#pragma unroll
      for (int c = 0; c < D; ++c) x[i * D + c] = m[i * D + c] + acc;
    }
#pragma unroll
    for (int e = 0; e < N * D; ++e) out[(size_t)(K - 1) * N * D + e] = x[e];
  }
  for (int k = K - 2; k >= 0; --k) {
    const size_t ci = inst * (size_t)(K - 1) + k;
    const double* A = a.cond_A + ci * N * N;
    const double* nm = a.cond_mean + ci * N * D;
    const double* nL = a.cond_chol + ci * N * N;
    const double* tl = a.cond_tl + ci * N;
    const double* to = a.cond_to + ci * N;
    const double* zk = z + (size_t)k * N;
    double xl[N * D], y[N * D];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int c = 0; c < D; ++c) xl[i * D + c] = tl[i] * x[i * D + c];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double toi = to[i];
      double lz = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) lz += (fabs(toi) * nL[i * N + j]) * zk[j];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) acc += (toi * A[i * N + j]) * xl[j * D + c];
        y[i * D + c] = (acc + nm[i * D + c]) + lz;
      }
    }
#pragma unroll
    for (int e = 0; e < N * D; ++e) {
      x[e] = y[e];
      out[(size_t)k * N * D + e] = y[e];
    }
  }
}

#ifndef PDQ_HOST_EMU
template <int N, int D>
__global__ void __launch_bounds__(64) lane_sample_kernel(SampleArgs a) {
  const size_t total = (size_t)a.n_inst * a.n_samples;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) lane_sample<N, D>(a, g);
}
#endif

}  // namespace pdq
