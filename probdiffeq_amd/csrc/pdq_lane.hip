// Instantiations and launcher of the lane-per-instance isotropic kernel.
#include <algorithm>
#include <atomic>

#include "pdq_lane.h"

namespace pdq {

constexpr int kMaxDevices = 64;

// The LDS attribute and the occupancy are set up once per device and shape, so
// that later launches (including one issued under stream capture) make no
// other runtime call than the launch itself.
template <int N, int D>
int lane_blocks_per_cu(int& per_cu) {
  static std::atomic<int> cache[kMaxDevices];
  int dev = 0;
  PDQ_HIP_TRY(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDevices && (per_cu = cache[dev].load()) > 0) return PDQ_OK;
  auto kern = lane_kernel<N, D>;
  const size_t lds = lane_lds_bytes<N, D>();
  PDQ_HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds));
  if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "lane kernel does not fit (%zu bytes of LDS)", lds);
  if (dev >= 0 && dev < kMaxDevices) cache[dev].store(per_cu);
  return PDQ_OK;
}

template <int N, int D>
int launch_lane_nd(LArgs& a, hipStream_t st) {
  auto kern = lane_kernel<N, D>;
  const size_t lds = lane_lds_bytes<N, D>();
  int per_cu = 0;
  const int rc = lane_blocks_per_cu<N, D>(per_cu);
  if (rc) return rc;
  const long need = ((long)a.n_inst + 63) / 64;
  const long cap = (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

bool lane_supported(int N, int D) { return D == 2 && (N == 4 || N == 5); }

int launch_lane(LArgs& a, int N, int D, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_lane_nd<4, 2>(a, st);
    if (N == 5) return launch_lane_nd<5, 2>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no lane kernel for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
