// Instantiations and launcher of the posterior sampler.
#include <algorithm>

#include "pdq_lane_sample.h"

namespace pdq {

namespace {

template <int N, int D>
int launch_sample_nd(SampleArgs& a, hipStream_t st) {
  const long total = (long)a.n_inst * a.n_samples;
  const long need = (total + 63) / 64;
  const long cap = 64L * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL((lane_sample_kernel<N, D>), dim3(grid), dim3(64), 0, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

}  // namespace

int launch_sample(SampleArgs& a, int N, int D, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_sample_nd<4, 2>(a, st);
    if (N == 5) return launch_sample_nd<5, 2>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no sampler for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
