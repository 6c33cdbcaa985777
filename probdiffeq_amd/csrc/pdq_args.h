// Kernel argument blocks and host-side launch entry points shared by the
// translation units of libpdq.so.
#pragma once

#ifdef PDQ_HOST_EMU
#ifndef PDQ_EMU_SHIM_INCLUDED
#include "pdq_hostemu.h"
#endif
#else
#include <hip/hip_runtime.h>
#endif

#include "../../include/pdq.h"
#include "pdq_layout.h"

namespace pdq {

struct KArgs {
  pdq_config cfg;
  Layout lay;
  int n_inst;
  int K;              // save_at / grid points
  int mode;           // MODE_SAVE_AT, MODE_FIXED_GRID, MODE_LML
  int every;          // MODE_LML: checkpoint stride on the grid
  int n_tan;          // MODE_LML: tangent directions (0: no gradient)
  int average_pdfs;
  int n_std;
  const double* params;
  const double* tcoeffs;  // or y0 in MODE_LML
  const double* times;    // save_at / grid (n_inst x K); MODE_LML: shared grid (K)
  const double* data;
  const double* stds;
  const double* A;        // nc x nc IWP transition (latent)
  const double* Q;        // nc x nc IWP noise factor (latent)
  pdq_solution out;
  double* lml_value;
  double* lml_grad;
  double* scratch;        // MODE_LML: per-instance checkpoint storage
  int* queue;             // work-queue counter (zeroed before launch)
};

enum { MODE_SAVE_AT = 0, MODE_FIXED_GRID = 1, MODE_LML = 2 };

// Arguments of the lane-per-instance isotropic kernel (pdq_lane.h).
struct LArgs {
  pdq_config cfg;
  int n_inst;
  int K;     // save_at / grid points per instance
  int mode;  // LMODE_SAVE_AT or LMODE_FIXED_GRID
  const double* params;   // n_inst x n_params
  const double* tcoeffs;  // n_inst x N x D
  const double* times;    // n_inst x K
  const double* A;        // N x N latent IWP transition
  const double* Q;        // N x N latent IWP noise factor
  pdq_solution out;
};
enum { LMODE_SAVE_AT = 0, LMODE_FIXED_GRID = 1 };

// Host helpers (pdq_capi.hip).
int host_error(int code, const char* fmt, ...);
int num_cus();

// Launchers (one per translation unit, compiled in parallel).
int launch_coop(KArgs& a, int G, hipStream_t st);
// Lane-per-instance isotropic kernel (pdq_lane.hip): shapes (N = nu + 1, d)
// it is instantiated for and its launcher.
bool lane_supported(int N, int D);
int launch_lane(LArgs& a, int N, int D, hipStream_t st);

// Lane-per-instance LML (pdq_lane_lml.hip).
// Arguments of the lane-per-instance LML kernel (pdq_lane_lml.h).
struct LmlArgs {
  pdq_config cfg;
  int n_inst, K, every, n_tan, average_pdfs, G;
  const double* params;
  const double* tcoeffs;  // n_inst x (N d) x (1 + n_tan): Taylor coefficients and their tangents
  const double* times;    // K grid points, shared by every instance
  const double* data;     // n_inst x Kobs x d
  const double* stds;     // n_inst x Kobs (one scalar per node)
  const double* Q;
  double* value;    // n_inst, or null
  double* grad;     // n_inst x n_tan, or null
  double* scratch;  // (Kobs - 1) x CondLayout<N, MC>::size x (n_inst G)
};

int lane_lml_tangents_per_lane(int n_tan);
bool lane_lml_supported(int N, int D, int solver);
size_t lane_lml_cond_size(int N, int D, int tpl);
int launch_lane_lml(LmlArgs& a, int N, int D, int tpl, hipStream_t st);
// The reverse-mode adjoint (pdq_lane_adjoint.h): one lane per instance; the
// scratch buffer holds lane_adj_scratch_doubles(...) doubles.
size_t lane_adj_scratch_doubles(int N, int D, int K, int every, int n_inst);
int launch_lane_adj(LmlArgs& a, int N, int D, hipStream_t st);
// Off-grid marginals (pdq_lane_offgrid.hip), isotropic shapes of the lane
// kernel, filter and fixed-interval smoother.
struct OffgridArgs {
  pdq_config cfg;
  int n_inst, K, n_query;
  const double* times;         // n_inst x K
  const double* t_query;       // n_inst x n_query
  const double* mean;          // the solution's marginals: n_inst x K x (N d)
  const double* chol;          // n_inst x K x N x N
  const double* filt_mean;     // fixed-interval smoother: filtering marginals, same layouts
  const double* filt_chol;
  const double* output_scale;  // n_inst x K
  const double* Q;
  double* mean_out;  // n_inst x n_query x (N d)
  double* chol_out;  // n_inst x n_query x N x N
};
int launch_offgrid(OffgridArgs& a, int N, int D, hipStream_t st);

// Off-grid marginals on the cooperative kernel's primitives (pdq_coop.hip):
// the dense factorisation, and isotropic shapes without a lane kernel.  k
// carries the solve's configuration, layout, grid (times) and transition
// matrices; the solution's arrays are in pdq_solution's layouts.
struct OffgridCoopArgs {
  KArgs k;
  int n_query;
  const double* t_query;       // n_inst x n_query
  const double* mean;          // n_inst x K x mean_size
  const double* chol;          // n_inst x K x chol_size
  const double* filt_mean;     // fixed-interval smoother: filtering marginals
  const double* filt_chol;
  const double* output_scale;  // n_inst x K
  double* mean_out;            // n_inst x n_query x mean_size
  double* chol_out;            // n_inst x n_query x chol_size
};
int launch_offgrid_coop(OffgridCoopArgs& a, int G, hipStream_t st);

// Posterior samples of a smoother (pdq_lane_sample.hip), isotropic.
struct SampleArgs {
  pdq_config cfg;
  int n_inst, K, n_samples;
  const double* term_mean;  // n_inst x (N d)
  const double* term_chol;  // n_inst x N x N
  const double* cond_A;     // n_inst x (K - 1) x N x N
  const double* cond_mean;  // n_inst x (K - 1) x (N d)
  const double* cond_chol;  // n_inst x (K - 1) x N x N
  const double* cond_tl;    // n_inst x (K - 1) x N
  const double* cond_to;    // n_inst x (K - 1) x N
  const double* base;       // n_inst x n_samples x K x N standard normal draws
  double* samples;          // n_inst x n_samples x K x (N d)
};
int launch_sample(SampleArgs& a, int N, int D, hipStream_t st);
// The same for any dense or isotropic shape with n * mc <= 32
// (pdq_sample_generic.h, pdq_post.hip); base has n draws per step.
constexpr int kSampleGenericMax = 32;  // n * mc
int launch_sample_generic(SampleArgs& a, int n, int mc, int dense, hipStream_t st);

// loss_lml_terminal_values on isotropic and dense marginals (pdq_post.hip);
// the dense kernel keeps a lane's (n + d) x d stack in LDS.
constexpr int kTerminalDenseBlock = 32;
constexpr int kTerminalDenseMax = 320;  // (n + d) d
int launch_terminal_lml(const pdq_config& cfg, int n_inst, int idx, const double* mean, const double* chol,
                        const double* u, const double* std, double* value, hipStream_t st);

int launch_taylor_tangents(int vf, int d, int np, int n_inst, int num, int ntan, const double* params,
                           const double* y0, const double* t0, double* out, hipStream_t st);
size_t coop_lds_bytes(const Layout& L, int G);

}  // namespace pdq

// A failed runtime call is reported, and the thread's last HIP error is
// cleared so that the caller's next HIP call does not pick it up.
#define PDQ_HIP_TRY(expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      (void)hipGetLastError();                                                             \
      return ::pdq::host_error(PDQ_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    }                                                                                      \
  } while (0)
