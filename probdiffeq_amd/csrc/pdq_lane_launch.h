// Launcher of one lane-kernel instantiation (shape N x D, strategy, solver).
// Each strategy's instantiations are compiled in a translation unit of their
// own (pdq_lane_<strategy>.hip) so that the units build in parallel.
#pragma once

#include <algorithm>
#include <atomic>

#include "pdq_lane.h"

namespace pdq {

constexpr int kLaneMaxDevices = 64;

template <int N, int D, int S, int V, bool O>
int launch_lane_nd(LArgs& a, hipStream_t st);

// Whether a configuration holds the defaults that the instantiations
// without OPTS fold in.
inline bool lane_default_options(const pdq_config& c) {
  return c.constraint == PDQ_TS0 && c.error_estimator == PDQ_ERROR_RESIDUAL_STD && !c.error_re_linearize &&
         !c.error_per_unit_step && !c.re_linearize_after_calibration && !c.init_std_rows;
}

#ifdef PDQ_LANE_DEFINE_LAUNCH
// The LDS attribute and the occupancy are set up once per device and
// instantiation, so that later launches (including one issued under stream
// capture) make no other runtime call than the launch itself.
template <int N, int D, int S, int V, bool O>
int lane_blocks_per_cu(int& per_cu) {
  static std::atomic<int> cache[kLaneMaxDevices];
  int dev = 0;
  PDQ_HIP_TRY(hipGetDevice(&dev));
  if (dev >= 0 && dev < kLaneMaxDevices && (per_cu = cache[dev].load()) > 0) return PDQ_OK;
  auto kern = lane_kernel<N, D, S, V, O>;
  const size_t lds = lane_lds_bytes<N, D, S>();
  PDQ_HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, lane_block_threads<S>(), lds));
  if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "lane kernel does not fit (%zu bytes of LDS)", lds);
  if (dev >= 0 && dev < kLaneMaxDevices) cache[dev].store(per_cu);
  return PDQ_OK;
}

template <int N, int D, int S, int V, bool O>
int launch_lane_nd(LArgs& a, hipStream_t st) {
  auto kern = lane_kernel<N, D, S, V, O>;
  const size_t lds = lane_lds_bytes<N, D, S>();
  int per_cu = 0;
  const int rc = lane_blocks_per_cu<N, D, S, V, O>(per_cu);
  if (rc) return rc;
  const long need = ((long)a.n_inst + 63) / 64;
  if (lane_split_strategy(S) && need > 0x7fffffffL) return host_error(PDQ_ERR_UNSUPPORTED, "too many instances");
  // the smoothers' kernels give every instance a lane of its own (no grid
  // stride); the filter's strides over at most the resident workgroups
  const long cap = lane_split_strategy(S) ? need : (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(lane_block_threads<S>()), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

// The shapes the device library instantiates, for one strategy.
#define PDQ_LANE_INSTANTIATE(S, O)                                                   \
  template int launch_lane_nd<4, 2, S, PDQ_SOLVER, O>(LArgs&, hipStream_t);         \
  template int launch_lane_nd<4, 2, S, PDQ_SOLVER_DYNAMIC, O>(LArgs&, hipStream_t); \
  template int launch_lane_nd<4, 2, S, PDQ_SOLVER_MLE, O>(LArgs&, hipStream_t);     \
  template int launch_lane_nd<5, 2, S, PDQ_SOLVER, O>(LArgs&, hipStream_t);         \
  template int launch_lane_nd<5, 2, S, PDQ_SOLVER_DYNAMIC, O>(LArgs&, hipStream_t); \
  template int launch_lane_nd<5, 2, S, PDQ_SOLVER_MLE, O>(LArgs&, hipStream_t);
#endif

}  // namespace pdq
