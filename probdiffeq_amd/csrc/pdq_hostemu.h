// Host build of the device sources (test infrastructure): compiled with g++
// and PDQ_HOST_EMU defined, the kernels' per-instance code runs on the CPU
// with one lane per group (G = 1) and one thread per workgroup, so that the
// kernel logic can be checked against the oracle without a GPU.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__

struct PdqEmuDim {
  unsigned x = 0;
};
static PdqEmuDim threadIdx;
static PdqEmuDim blockDim{1};

inline void __syncthreads() {}
inline void __threadfence_block() {}
inline double __shfl_xor(double v, int, int) { return v; }
inline int __shfl(int v, int, int) { return v; }
inline double __shfl(double v, int, int) { return v; }
inline int atomicAdd(int* p, int v) {
  const int old = *p;
  *p += v;
  return old;
}
#define __builtin_amdgcn_fence(a, b) ((void)0)
#define __builtin_amdgcn_wave_barrier() ((void)0)

using std::copysign;
using std::fabs;
using std::fmax;
using std::fmin;
using std::hypot;
using std::isfinite;
using std::log;
using std::pow;
using std::sqrt;
typedef void* hipStream_t;
