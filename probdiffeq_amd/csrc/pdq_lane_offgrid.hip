// Instantiations and launcher of the off-grid marginal kernel.
#include <algorithm>

#include "pdq_lane_offgrid.h"

namespace pdq {

namespace {

template <int N, int D, int S>
int launch_offgrid_nd(OffgridArgs& a, hipStream_t st) {
  const long total = (long)a.n_inst * a.n_query;
  const long need = (total + 63) / 64;
  const long cap = 64L * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL((lane_offgrid_kernel<N, D, S>), dim3(grid), dim3(64), 0, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

template <int N, int D>
int launch_offgrid_shape(OffgridArgs& a, hipStream_t st) {
  if (a.cfg.strategy == PDQ_FILTER) return launch_offgrid_nd<N, D, PDQ_FILTER>(a, st);
  return launch_offgrid_nd<N, D, PDQ_FIXEDINTERVAL>(a, st);
}

}  // namespace

int launch_offgrid(OffgridArgs& a, int N, int D, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_offgrid_shape<4, 2>(a, st);
    if (N == 5) return launch_offgrid_shape<5, 2>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no off-grid marginal kernel for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
