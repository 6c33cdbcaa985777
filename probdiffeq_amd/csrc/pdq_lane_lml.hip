// Instantiations and launcher of the lane-per-instance LML kernel.
#include <algorithm>
#include <atomic>

#include "pdq_lane_adjoint.h"
#include "pdq_lane_lml.h"

namespace pdq {

namespace {

constexpr int kLmlMaxDevices = 64;

template <int N, int D, int TPL, int S>
int launch_lml_nd(LmlArgs& a, hipStream_t st) {
  constexpr int MC = D * (1 + TPL);
  // the forward pass's merges on a second wavefront where three slots with a
  // tangent column fit (decided per shape, so that the value-only and the
  // gradient launches of one shape run the same form)
  constexpr bool split = lane_lml_split_fits<N, D * 2>();
  auto kern = lane_lml_kernel<N, D, TPL, S, split>;
  const size_t lds = lane_lml_lds_bytes<N, MC, split>();
  // LDS attribute and occupancy once per device and instantiation
  static std::atomic<int> cache[kLmlMaxDevices];
  int dev = 0;
  PDQ_HIP_TRY(hipGetDevice(&dev));
  int per_cu = (dev >= 0 && dev < kLmlMaxDevices) ? cache[dev].load() : 0;
  if (per_cu <= 0) {
    PDQ_HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, split ? 128 : 64, lds));
    if (per_cu < 1) return host_error(PDQ_ERR_UNSUPPORTED, "lane LML kernel does not fit (%zu bytes of LDS)", lds);
    if (dev >= 0 && dev < kLmlMaxDevices) cache[dev].store(per_cu);
  }
  const long lanes = (long)a.n_inst * a.G;
  const long need = (lanes + 63) / 64;
  if (need > 0x7fffffffL) return host_error(PDQ_ERR_UNSUPPORTED, "too many instances");
  // SPLIT: one lane of wavefront 0 per (instance, tangent group), no stride
  const long cap = split ? need : (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min(need, cap));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(split ? 128 : 64), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

template <int N, int D, int TPL>
int launch_lml_tpl(LmlArgs& a, hipStream_t st) {
  if (a.cfg.solver == PDQ_SOLVER_DYNAMIC) return launch_lml_nd<N, D, TPL, PDQ_SOLVER_DYNAMIC>(a, st);
  return launch_lml_nd<N, D, TPL, PDQ_SOLVER>(a, st);
}

template <int N, int D>
int launch_lml_shape(LmlArgs& a, int tpl, hipStream_t st) {
  switch (tpl) {
    case 0: return launch_lml_tpl<N, D, 0>(a, st);
    case 1: return launch_lml_tpl<N, D, 1>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no lane LML kernel with %d tangents per lane", tpl);
}

template <int N, int D, int S>
int launch_adj_nd(LmlArgs& a, hipStream_t st) {
  // two wavefronts per workgroup (the segment recomputation beside the
  // backward walk) where their four slots per lane fit the LDS
  constexpr bool split = lane_adj_lds_bytes<N, D, true>() <= 160 * 1024;
  auto kern = lane_adj_kernel<N, D, S, split>;
  const size_t lds = lane_adj_lds_bytes<N, D, split>();
  static std::atomic<int> done[kLmlMaxDevices];
  int dev = 0;
  PDQ_HIP_TRY(hipGetDevice(&dev));
  if (!(dev >= 0 && dev < kLmlMaxDevices && done[dev].load())) {
    PDQ_HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (dev >= 0 && dev < kLmlMaxDevices) done[dev].store(1);
  }
  // one lane per instance: the grid covers every instance once (the scratch
  // buffer has a column per lane of wavefront 0 of this grid)
  const int grid = (a.n_inst + 63) / 64;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(split ? 128 : 64), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

}  // namespace

size_t lane_adj_scratch_doubles(int N, int D, int K, int every, int n_inst) {
  const size_t lanes = (size_t)((n_inst + 63) / 64) * 64;
  if (D == 2 && N == 4) return AdjLayout<4, 2>(K, every).size * lanes;
  if (D == 2 && N == 5) return AdjLayout<5, 2>(K, every).size * lanes;
  return 0;
}

int launch_lane_adj(LmlArgs& a, int N, int D, hipStream_t st) {
  if (D == 2 && (N == 4 || N == 5)) {
    const bool dyn = a.cfg.solver == PDQ_SOLVER_DYNAMIC;
    if (N == 4) return dyn ? launch_adj_nd<4, 2, PDQ_SOLVER_DYNAMIC>(a, st) : launch_adj_nd<4, 2, PDQ_SOLVER>(a, st);
    return dyn ? launch_adj_nd<5, 2, PDQ_SOLVER_DYNAMIC>(a, st) : launch_adj_nd<5, 2, PDQ_SOLVER>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no adjoint LML kernel for nu + 1 = %d, d = %d", N, D);
}

// One tangent column per lane: G = n_tan lanes per instance.  Against three
// per lane (G = 2) this is 1.44x faster on the C5 batch (13.9 vs 20.0 ms):
// the extra lanes are free (the batch occupies a fraction of the CUs) and each
// lane's mean recursions shrink from 8 to 4 columns.
int lane_lml_tangents_per_lane(int n_tan) { return n_tan == 0 ? 0 : 1; }

bool lane_lml_supported(int N, int D, int solver) {
  return D == 2 && (N == 4 || N == 5) && (solver == PDQ_SOLVER || solver == PDQ_SOLVER_DYNAMIC);
}

size_t lane_lml_cond_size(int N, int D, int tpl) {
  const size_t mc = (size_t)D * (1 + tpl);
  return (size_t)N * N + (size_t)N * mc + (size_t)N * (N + 1) / 2 + 2 * (size_t)N;
}

int launch_lane_lml(LmlArgs& a, int N, int D, int tpl, hipStream_t st) {
  if (D == 2) {
    if (N == 4) return launch_lml_shape<4, 2>(a, tpl, st);
    if (N == 5) return launch_lml_shape<5, 2>(a, tpl, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "no lane LML kernel for nu + 1 = %d, d = %d", N, D);
}

}  // namespace pdq
