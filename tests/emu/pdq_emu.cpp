// CPU build of the kernels' per-instance code (test infrastructure).
// Compiled with g++ -DPDQ_HOST_EMU from the product's device sources; runs
// every instance with one lane (G = 1) so that the kernel logic can be
// compared with the oracle on the CPU.
#include <cstring>
#include <vector>

#include "../../probdiffeq_amd/csrc/pdq_blockdiag.hip"
#include "../../probdiffeq_amd/csrc/pdq_coop.hip"
#include "../../probdiffeq_amd/csrc/pdq_iwp.h"

namespace pdq {
int host_error(int code, const char*, ...) { return code; }
int num_cus() { return 1; }
}  // namespace pdq

extern "C" {

// mode: 0 save_at, 1 fixed grid.  Output pointers as in pdq_solution (host).
int pdqemu_solve(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs, const double* times,
                 int K, const pdq_solution* out, int mode) {
  int queue = 0;
  if (cfg->factorisation == PDQ_BLOCKDIAG) {
    std::vector<double> A1, Q1;
    pdq::iwp_matrices(cfg->num_derivatives, A1, Q1);
    pdq::BArgs a{};
    a.cfg = *cfg;
    a.n_inst = n_inst;
    a.K = K;
    a.mode = mode;
    a.params = params;
    a.tcoeffs = tcoeffs;
    a.times = times;
    a.A1 = A1.data();
    a.Q1 = Q1.data();
    a.out = *out;
    a.queue = &queue;
    std::vector<double> y(cfg->ode_dim), red(16);
    const int N = cfg->num_derivatives + 1, d = cfg->ode_dim;
    for (int i = 0; i < n_inst; ++i) {
#define PDQ_EMU_BD(NN, BB)                                               \
  if (N == NN && d <= BB) {                                              \
    pdq::emu_solve_instance_bd<NN, BB>(a, i, y.data(), red.data());       \
    continue;                                                            \
  }
      PDQ_EMU_BD(3, 2)
      PDQ_EMU_BD(3, 64)
      PDQ_EMU_BD(5, 2)
      PDQ_EMU_BD(4, 2)
      return 2;
    }
    return 0;
  }
  pdq::Layout L = pdq::make_layout(*cfg, 0);
  std::vector<double> A, Q;
  if (cfg->factorisation == PDQ_DENSE) {
    pdq::iwp_dense(cfg->num_derivatives, cfg->ode_dim, cfg->base_output_scale, A, Q);
  } else {
    pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  }
  pdq::KArgs a{};
  a.cfg = *cfg;
  a.lay = L;
  a.n_inst = n_inst;
  a.K = K;
  a.mode = mode;
  a.params = params;
  a.tcoeffs = tcoeffs;
  a.times = times;
  a.A = A.data();
  a.Q = Q.data();
  a.out = *out;
  a.queue = &queue;
  std::vector<double> shared(L.wg_shared), ws(L.inst_size);
  for (int i = 0; i < n_inst; ++i) pdq::emu_solve_instance(a, i, shared.data(), ws.data());
  return 0;
}

}  // extern "C"
