// Instantiations and launcher of the lane-per-instance isotropic kernel.
#include <algorithm>

#include "pdq_lane.h"

namespace pdq {

template <int N, int D>
int launch_lane_nd(LArgs& a, hipStream_t st) {
  auto kern = lane_kernel<N, D>;
  int per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
  if (per_cu < 1) per_cu = 1;
  const long need = ((long)a.n_inst + 63) / 64;
  const long cap = (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

bool lane_supported(int N, int D) { return D == 2 && (N == 4 || N == 5); }

size_t lane_workspace_doubles(int N, int D, int n_inst) { return lane_ws_doubles(N, D, n_inst); }

int launch_lane(LArgs& a, int N, int D, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_lane_nd<4, 2>(a, st);
    if (N == 5) return launch_lane_nd<5, 2>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no lane kernel for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
