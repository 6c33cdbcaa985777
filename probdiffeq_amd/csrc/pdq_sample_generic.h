// Samples of a smoother's posterior (MarkovSequence.sample,
// estimators_and_losses.py:233-273) for the shapes without a lane sampler:
// the dense factorisation and isotropic shapes other than nu = 3, 4 with
// d = 2, one lane per (instance, sample), n * mc <= 32 (n the covariance
// dimension, mc the mean columns).
//
// The reverse scan starts from a draw of the terminal marginal and pushes the
// sample back through the backward conditionals: predicted =
// cond.apply_flat(x), x = predicted.sample_flat(z).
// * dense (ssm_impl_dense.py:18-22): mean = to * (A (tl * x) + m), factor
//   |to| * L, and an n-vector of draws per step;
// * isotropic (ssm_impl_isotropic.py:75-79, 255-258): mean = (to * A)
//   (tl * x) + m, without to on the noise mean, and one N-vector of draws that
//   the d columns share.
// The draws come in as data: base[instance, sample, k] (n entries) for
// conditional k, row K - 1 for the terminal marginal.  The lane's vectors
// live in `v` with element stride `stride` (LDS on the device).
#pragma once

#include "pdq_args.h"

namespace pdq {

#if defined(__HIPCC__)
__host__ __device__
#endif
inline void sample_generic(const SampleArgs& a, size_t gid, int n, int mc, int dense, double* v, int stride) {
  const int K = a.K, ms = n * mc;
  const size_t inst = gid / (size_t)a.n_samples;
  const double* z = a.base + gid * (size_t)K * n;
  double* out = a.samples + gid * (size_t)K * ms;
  double* x = v;
  double* xl = v + kSampleGenericMax * stride;
  double* y = v + 2 * kSampleGenericMax * stride;
  {
    const double* m = a.term_mean + inst * ms;
    const double* L = a.term_chol + inst * (size_t)n * n;
    const double* zk = z + (size_t)(K - 1) * n;
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc += L[i * n + j] * zk[j];
      for (int c = 0; c < mc; ++c) {
        x[(i * mc + c) * stride] = m[i * mc + c] + acc;
        out[(size_t)(K - 1) * ms + i * mc + c] = x[(i * mc + c) * stride];
      }
    }
  }
  for (int k = K - 2; k >= 0; --k) {
    const size_t ci = inst * (size_t)(K - 1) + k;
    const double* A = a.cond_A + ci * n * n;
    const double* nm = a.cond_mean + ci * ms;
    const double* nL = a.cond_chol + ci * n * n;
    const double* tl = a.cond_tl + ci * n;
    const double* to = a.cond_to + ci * n;
    const double* zk = z + (size_t)k * n;
    for (int e = 0; e < ms; ++e) xl[e * stride] = tl[e / mc] * x[e * stride];
    for (int i = 0; i < n; ++i) {
      const double toi = to[i];
      double lz = 0.0;
      for (int j = 0; j < n; ++j) lz += (fabs(toi) * nL[i * n + j]) * zk[j];
      for (int c = 0; c < mc; ++c) {
        double acc = 0.0;
        if (dense) {
          for (int j = 0; j < n; ++j) acc += A[i * n + j] * xl[(j * mc + c) * stride];
          y[(i * mc + c) * stride] = toi * (acc + nm[i * mc + c]) + lz;
        } else {
          for (int j = 0; j < n; ++j) acc += (toi * A[i * n + j]) * xl[(j * mc + c) * stride];
          y[(i * mc + c) * stride] = (acc + nm[i * mc + c]) + lz;
        }
      }
    }
    for (int e = 0; e < ms; ++e) {
      x[e * stride] = y[e * stride];
      out[(size_t)k * ms + e] = y[e * stride];
    }
  }
}

}  // namespace pdq
