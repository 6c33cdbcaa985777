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

}  // namespace

int launch_terminal_lml(const pdq_config& cfg, int n_inst, int idx, const double* mean, const double* chol,
                        const double* u, const double* std, double* value, hipStream_t st) {
  const int N = cfg.num_derivatives + 1, d = cfg.ode_dim;
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
