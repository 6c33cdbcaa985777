// Losses on a solution's marginals, one lane per instance.
//
// loss_lml_terminal_values (estimators_and_losses.py:20-50), isotropic:
// marginals.to_derivative(idx, std) observes row idx of the (N x d) mean with
// noise std; marginalise (ssm_impl_isotropic.py:81-89) stacks the factor's
// row idx over std and QR-reduces the (N + 1) x 1 column, so the observed
// factor is +-||(L[idx, :], std)||; logpdf_tree (:213-236) sums the scalar
// log-densities of the d data columns against that factor.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pdq_args.h"
#include "pdq_sample_generic.h"

namespace pdq {

namespace {

__global__ void __launch_bounds__(64) terminal_lml_kernel(int N, int d, int n_inst, int idx, const double* mean,
                                                          const double* chol, const double* u, const double* std,
                                                          double* value) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= n_inst) return;
  const double* L = chol + (size_t)inst * N * N + (size_t)idx * N;
  const double* m = mean + (size_t)inst * N * d + (size_t)idx * d;
  const double* y = u + (size_t)inst * d;
  const double s = std[inst];
  double ss = s * s;
  for (int j = N - 1; j >= 0; --j) ss = fma(L[j], L[j], ss);
  const double r = sqrt(ss);
  const double logdet = 2.0 * log(r);
  const double c2pi = log(M_PI * 2.0);
  double lp = 0.0;
  for (int k = 0; k < d; ++k) {
    const double w = (y[k] - m[k]) / r;
    lp += -0.5 * (logdet + w * w + 1.0 * c2pi);
  }
  value[inst] = lp;
}

// The dense form (ssm_impl_dense.py:24-33 marginalise, :170-180 logpdf):
// derivative idx's d coordinates observed with noise diag(std) (std: d
// entries per instance).  The observed factor is R^T with R the R-only QR
// of the stacked [(H L)^T; diag(std)] ((n + d) x d, LAPACK signs); logpdf
// re-triangularises it (R is already triangular, so it stays R), and
// log p = -|R^-T (u - m)|^2 / 2 - d log(2 pi) / 2 - sum log|R_ii|.  One lane
// per instance; its stack lives in LDS with stride blockDim.x.
__global__ void __launch_bounds__(kTerminalDenseBlock) terminal_lml_dense_kernel(int n, int d, int n_inst, int idx,
                                                                                  const double* mean,
                                                                                  const double* chol, const double* u,
                                                                                  const double* std, double* value) {
  extern __shared__ double lds[];
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= n_inst) return;  // no barriers below
  const int S = blockDim.x, rows = n + d;
  double* M = lds + threadIdx.x;  // M[r][c] at (r * d + c) * S
  const double* L = chol + (size_t)inst * n * n;
  const double* sd = std + (size_t)inst * d;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < d; ++c)
      M[(r * d + c) * S] = r < n ? L[(idx * d + c) * n + r] : (r - n == c ? sd[c] : 0.0);
  for (int i = 0; i < d; ++i) {
    double ss = 0.0;
    for (int r = i + 1; r < rows; ++r) ss += M[(r * d + i) * S] * M[(r * d + i) * S];
    if (ss == 0.0) continue;
    const double alpha = M[(i * d + i) * S];
    const double beta = -copysign(hypot(alpha, sqrt(ss)), alpha);
    const double tau = (beta - alpha) / beta, scal = 1.0 / (alpha - beta);
    for (int r = i + 1; r < rows; ++r) M[(r * d + i) * S] *= scal;
    M[(i * d + i) * S] = beta;
    for (int j = i + 1; j < d; ++j) {
      double w = M[(i * d + j) * S];
      for (int r = i + 1; r < rows; ++r) w += M[(r * d + i) * S] * M[(r * d + j) * S];
      w *= tau;
      M[(i * d + j) * S] -= w;
      for (int r = i + 1; r < rows; ++r) M[(r * d + j) * S] -= M[(r * d + i) * S] * w;
    }
  }
  const double* m = mean + (size_t)inst * n + (size_t)idx * d;
  const double* y = u + (size_t)inst * d;
  double slogdet = 0.0, sq = 0.0;
  for (int i = 0; i < d; ++i) slogdet += log(fabs(M[(i * d + i) * S]));
  // forward substitution with the lower factor R^T; w overwrites row n of M
  double* w = M + (size_t)n * d * S;
  for (int i = 0; i < d; ++i) {
    double s = y[i] - m[i];
    for (int j = 0; j < i; ++j) s -= M[(j * d + i) * S] * w[j * S];
    w[i * S] = s / M[(i * d + i) * S];
    sq += w[i * S] * w[i * S];
  }
  const double c2pi = log(2.0 * M_PI);
  value[inst] = -1.0 / 2.0 * sq - (double)d / 2.0 * c2pi - slogdet;
}

}  // namespace

int launch_terminal_lml(const pdq_config& cfg, int n_inst, int idx, const double* mean, const double* chol,
                        const double* u, const double* std, double* value, hipStream_t st) {
  const int N = cfg.num_derivatives + 1, d = cfg.ode_dim;
  if (cfg.factorisation == PDQ_DENSE) {
    const int n = N * d;
    if ((n + d) * d > kTerminalDenseMax)
      return host_error(PDQ_ERR_UNSUPPORTED, "loss_lml_terminal_values (dense) supports (n + d) d <= %d", kTerminalDenseMax);
    const size_t lds = sizeof(double) * (size_t)kTerminalDenseBlock * (n + d) * d;
    PDQ_HIP_TRY(hipFuncSetAttribute((const void*)terminal_lml_dense_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(terminal_lml_dense_kernel, dim3((n_inst + kTerminalDenseBlock - 1) / kTerminalDenseBlock),
                       dim3(kTerminalDenseBlock), lds, st, n, d, n_inst, idx, mean, chol, u, std, value);
    PDQ_HIP_TRY(hipGetLastError());
    return PDQ_OK;
  }
  hipLaunchKernelGGL(terminal_lml_kernel, dim3((n_inst + 63) / 64), dim3(64), 0, st, N, d, n_inst, idx, mean, chol, u,
                     std, value);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

// MarkovSequence.sample for the shapes without a lane sampler: one lane per
// (instance, sample), its three vectors element-major in LDS (stride 64, so
// the lanes of a wavefront hit distinct banks).
__global__ void __launch_bounds__(64) sample_generic_kernel(SampleArgs a, int n, int mc, int dense) {
  __shared__ double v[3 * kSampleGenericMax * 64];
  const size_t total = (size_t)a.n_inst * a.n_samples;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride)
    sample_generic(a, g, n, mc, dense, v + threadIdx.x, 64);
}

int launch_sample_generic(SampleArgs& a, int n, int mc, int dense, hipStream_t st) {
  if (n * mc > kSampleGenericMax) return host_error(PDQ_ERR_UNSUPPORTED, "sampling supports n * mc <= %d", kSampleGenericMax);
  const size_t total = (size_t)a.n_inst * a.n_samples;
  const long need = (long)((total + 63) / 64);
  const long cap = 8L * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(sample_generic_kernel, dim3(grid), dim3(64), 0, st, a, n, mc, dense);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

}  // namespace pdq
