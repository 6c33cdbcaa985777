// The cooperative kernel for any dense / isotropic shape (dimensions read at
// run time), the Taylor-tangent kernel of the LML path, and the dispatcher
// that prefers a compile-time-shape instantiation when one exists.
#include "pdq_coop.h"

namespace pdq {

// Initial Taylor coefficients with forward-mode tangents d/d params and
// d/d y0 (jet_expansion_algorithms.py:49-103 for the registry); one thread
// per instance.  out: n_inst x (N d) x (1 + ntan).
__global__ void taylor_tangent_kernel(int vf, int d, int np, int n_inst, int num, int ntan, const double* params,
                                      const double* y0, const double* t0p, double* out) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= n_inst) return;
  const double t0 = t0p[0];
  Dual prmD[kMaxParams];
  Dual cc[64];
  for (int q = 0; q < np; ++q) {
    prmD[q] = dconst(params[(size_t)inst * np + q]);
    if (q < ntan) prmD[q].g[q] = 1.0;
  }
  for (int k = 0; k < d; ++k) {
    cc[k] = dconst(y0[(size_t)inst * d + k]);
    if (np + k < ntan) cc[k].g[np + k] = 1.0;
  }
  for (int k = 0; k < num; ++k)
    for (int j = 0; j < d; ++j) cc[(k + 1) * d + j] = taylor_rhs(vf, d, cc, prmD, t0, k, j) / (double)(k + 1);
  const int N = num + 1;
  for (int i = 0; i < N; ++i) {
    const double f = factorial_d(i);
    for (int k = 0; k < d; ++k) {
      const Dual v = f * cc[i * d + k];
      for (int q = 0; q <= ntan; ++q)
        out[(((size_t)inst * N + i) * d + k) * (1 + ntan) + q] = q == 0 ? v.v : v.g[q - 1];
    }
  }
}

int launch_taylor_tangents(int vf, int d, int np, int n_inst, int num, int ntan, const double* params,
                           const double* y0, const double* t0, double* out, hipStream_t st) {
  hipLaunchKernelGGL(taylor_tangent_kernel, dim3((n_inst + 63) / 64), dim3(64), 0, st, vf, d, np, n_inst, num, ntan,
                     params, y0, t0, out);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

int launch_coop_fixed(KArgs& a, int G, hipStream_t st) {
  const Layout& L = a.lay;
  const int dense = a.cfg.factorisation == PDQ_DENSE;
#define PDQ_MATCH_SHAPE(name, NC, MC, MCP, KO, N_, D_, DENSE_)                                          \
  if (L.nc == NC && L.mc == MC && L.mcp == MCP && L.ko == KO && L.N == N_ && L.d == D_ && dense == DENSE_) \
    return launch_coop_##name(a, G, st);
  PDQ_FIXED_SHAPES(PDQ_MATCH_SHAPE)
#undef PDQ_MATCH_SHAPE
  return kNoFixedShape;
}

size_t coop_lds_bytes(const Layout& L, int G) { return sizeof(double) * (size_t)(L.wg_shared + (64 / G) * L.inst_size); }

int launch_coop(KArgs& a, int G, hipStream_t st) {
  const int rc = launch_coop_fixed(a, G, st);
  if (rc != kNoFixedShape) return rc;
  switch (G) {
    case 4: return launch_coop_g<4, DynDims>(a, st);
    case 8: return launch_coop_g<8, DynDims>(a, st);
    case 16: return launch_coop_g<16, DynDims>(a, st);
    case 32: return launch_coop_g<32, DynDims>(a, st);
    case 64: return launch_coop_g<64, DynDims>(a, st);
  }
  return host_error(PDQ_ERR_INVALID, "lanes_per_instance must be one of 4, 8, 16, 32, 64 (got %d)", G);
}

}  // namespace pdq
