// Launcher of one block-diagonal kernel instantiation (block size N x N,
// strategy).  Each strategy's instantiations are compiled in a translation
// unit of their own (pdq_bd_<strategy>.hip) so that the units build in
// parallel.
#pragma once

#include <algorithm>
#include <atomic>

#include "pdq_bd.h"

namespace pdq {

template <int N, int S>
int launch_bd_ns(BdArgs& a, hipStream_t st);

#ifdef PDQ_BD_DEFINE_LAUNCH
// Resident workgroups per CU of a kernel, queried once per device, so that a
// later launch makes no other runtime call than the launch itself.
template <int N, int S, bool MULTI>
int bd_blocks_per_cu(int threads, int& per_cu) {
  constexpr int kMaxDev = 64;
  static std::atomic<int> cache[kMaxDev][kBdMaxDim / 64 + 1];
  int dev = 0;
  PDQ_HIP_TRY(hipGetDevice(&dev));
  const int slot = threads / 64;
  if (dev >= 0 && dev < kMaxDev && (per_cu = cache[dev][slot].load()) > 0) return PDQ_OK;
  per_cu = 0;
  auto kern = MULTI ? bd_block_kernel<N, S> : bd_wave_kernel<N, S>;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
  if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "the block-diagonal kernel does not fit (%d lanes)", threads);
  if (dev >= 0 && dev < kMaxDev) cache[dev][slot].store(per_cu);
  return PDQ_OK;
}

template <int N, int S>
int launch_bd_ns(BdArgs& a, hipStream_t st) {
  int per_cu = 0;
  if (a.G <= 64) {
    const int rc = bd_blocks_per_cu<N, S, false>(64, per_cu);
    if (rc) return rc;
    const int gpw = 64 / a.G;
    const long need = ((long)a.n_inst + gpw - 1) / gpw;
    const int grid = (int)std::max<long>(1, std::min(need, (long)per_cu * num_cus()));
    auto kern = bd_wave_kernel<N, S>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, st, a);
  } else {
    const int rc = bd_blocks_per_cu<N, S, true>(a.G, per_cu);
    if (rc) return rc;
    const int grid = (int)std::max<long>(1, std::min((long)a.n_inst, (long)per_cu * num_cus()));
    auto kern = bd_block_kernel<N, S>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(a.G), 0, st, a);
  }
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

#define PDQ_BD_INSTANTIATE(S)                                   \
  template int launch_bd_ns<2, S>(BdArgs&, hipStream_t);       \
  template int launch_bd_ns<3, S>(BdArgs&, hipStream_t);       \
  template int launch_bd_ns<4, S>(BdArgs&, hipStream_t);       \
  template int launch_bd_ns<5, S>(BdArgs&, hipStream_t);       \
  template int launch_bd_ns<6, S>(BdArgs&, hipStream_t);
#endif

}  // namespace pdq
