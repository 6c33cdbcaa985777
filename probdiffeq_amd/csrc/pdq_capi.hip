// C ABI of the gfx950 solver path (declared in include/pdq.h).
//
// Host responsibilities: validate the configuration, build the IWP system
// matrices once per (nu, d, factorisation, base scale), size the LDS layout,
// pick lanes-per-instance (autotuned or heuristic), zero the work-queue
// counter with a kernel (no memset node, so launches stay capturable) and
// launch the persistent kernels.  No compute ever falls back to the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "pdq_args.h"
#include "pdq_bd.h"
#include "pdq_device.h"
#include "pdq_iwp.h"

namespace {

// The last error message of the calling thread (a fixed buffer: setting it
// allocates nothing).
thread_local char g_err[512];

}  // namespace

namespace pdq {
int host_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  std::memcpy(g_err, buf, sizeof(g_err));
  return code;
}

int num_cus() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
  return v;
}
}  // namespace pdq

namespace {

int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// A failed runtime call is reported, and the thread's last HIP error is
// cleared so that the caller's next HIP call does not pick it up.
#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      (void)hipGetLastError();                                                   \
      return set_err(PDQ_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    }                                                                            \
  } while (0)

struct DevMats {
  double* A = nullptr;
  double* Q = nullptr;
};

// Stream-ordered device buffers of one call, freed (stream ordered) on every
// exit path of the call.
struct CallBuffers {
  hipStream_t st;
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  int n = 0;
  explicit CallBuffers(hipStream_t s) : st(s) {}
  CallBuffers(const CallBuffers&) = delete;
  CallBuffers& operator=(const CallBuffers&) = delete;
  ~CallBuffers() {
    for (int i = n - 1; i >= 0; --i) (void)hipFreeAsync(p[i], st);
  }
  template <class T>
  hipError_t alloc(T** out, size_t bytes) {
    void* q = nullptr;
    const hipError_t e = hipMallocAsync(&q, bytes, st);
    if (e != hipSuccess) return e;
    p[n++] = q;
    *out = static_cast<T*>(q);
    return hipSuccess;
  }
};

std::mutex g_mu;
std::map<std::tuple<int, int, int, int, double, int>, DevMats> g_mats;  // (dev, fact, nu, d, base, kind)
// (mode, factorisation, nu, d, strategy, solver, constraint, vector field,
// prior, batch-size bucket): a POD key, so that a lookup allocates nothing
using TuneKey = std::array<int, 10>;
std::map<TuneKey, int> g_tuned;

int device_mats(const pdq_config& cfg, bool latent_1d, DevMats& out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, latent_1d ? -1 : cfg.factorisation, cfg.num_derivatives,
                                   latent_1d ? 0 : cfg.ode_dim, latent_1d ? 0.0 : cfg.base_output_scale, 0);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mats.find(key);
  if (it != g_mats.end()) {
    out = it->second;
    return PDQ_OK;
  }
  std::vector<double> A1, Q1;
  pdq::iwp_matrices(cfg.num_derivatives, A1, Q1);
  const int N = cfg.num_derivatives + 1, d = cfg.ode_dim;
  std::vector<double> A, Q;
  if (!latent_1d && cfg.factorisation == PDQ_DENSE) {
    const int n = N * d;
    A.assign(n * n, 0.0);
    Q.assign(n * n, 0.0);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        for (int k = 0; k < d; ++k) {
          A[(i * d + k) * n + j * d + k] = A1[i * N + j];
          Q[(i * d + k) * n + j * d + k] = Q1[i * N + j] * cfg.base_output_scale;  // kron(Q1, Lambda)
        }
  } else {
    A = A1;
    Q = Q1;
  }
  DevMats m;
  HIP_TRY(hipMalloc(&m.A, A.size() * sizeof(double)));
  HIP_TRY(hipMalloc(&m.Q, Q.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(m.A, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m.Q, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
  g_mats[key] = m;
  out = m;
  return PDQ_OK;
}

// The registry's requirements on a vector field's dimension and parameters
// (pdq_device.h reads params[0 .. n_params) and y[0 .. d) only within them).
int check_vector_field(int vf, int d, int n_params) {
  if (vf < 0 || vf > PDQ_VF_LOGISTIC) return set_err(PDQ_ERR_UNSUPPORTED, "unknown vector field %d", vf);
  if (d < 1) return set_err(PDQ_ERR_INVALID, "ode_dim must be positive");
  if (n_params < 0 || n_params > pdq::kMaxParams) return set_err(PDQ_ERR_INVALID, "n_params out of range");
  if ((vf == PDQ_VF_LOTKA_VOLTERRA || vf == PDQ_VF_VAN_DER_POL) && d != 2)
    return set_err(PDQ_ERR_INVALID, "this vector field needs ode_dim == 2");
  if (vf == PDQ_VF_LOTKA_VOLTERRA && n_params < 4) return set_err(PDQ_ERR_INVALID, "lotka_volterra needs 4 params");
  if (vf == PDQ_VF_VAN_DER_POL && n_params < 1) return set_err(PDQ_ERR_INVALID, "van_der_pol needs 1 param");
  if (vf == PDQ_VF_BURGERS && n_params < 2) return set_err(PDQ_ERR_INVALID, "burgers needs (viscosity, dx)");
  if (vf == PDQ_VF_LINEAR && n_params < d) return set_err(PDQ_ERR_INVALID, "linear needs d params");
  if (vf == PDQ_VF_LOGISTIC && n_params < 1) return set_err(PDQ_ERR_INVALID, "logistic needs 1 param");
  return PDQ_OK;
}

int check_device() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(PDQ_ERR_NO_DEVICE, "no HIP device");
  return PDQ_OK;
}

// The configuration checks shared by the entry points; with `device` the
// presence of a HIP device is checked last.
int check_common(const pdq_config* cfg, int n_inst, int K, bool device = true) {
  int rc = PDQ_OK;
  if (!cfg) return set_err(PDQ_ERR_INVALID, "cfg is NULL");
  if (n_inst < 0) return set_err(PDQ_ERR_INVALID, "n_inst < 0");
  if (K < 2) return set_err(PDQ_ERR_INVALID, "need at least two time points (got %d)", K);
  if (cfg->num_derivatives < 1 || cfg->num_derivatives > 7)
    return set_err(PDQ_ERR_UNSUPPORTED, "num_derivatives must be in [1, 7] (got %d)", cfg->num_derivatives);
  if (cfg->ode_dim < 1) return set_err(PDQ_ERR_INVALID, "ode_dim must be positive");
  rc = check_vector_field(cfg->vector_field, cfg->ode_dim, cfg->n_params);
  if (rc) return rc;
  if (cfg->strategy < 0 || cfg->strategy > 2 || cfg->solver < 0 || cfg->solver > 2 || cfg->constraint < 0 ||
      cfg->constraint > 1 || cfg->factorisation < 0 || cfg->factorisation > 2)
    return set_err(PDQ_ERR_INVALID, "invalid enum value in configuration");
  if (cfg->error_estimator != PDQ_ERROR_RESIDUAL_STD && cfg->error_estimator != PDQ_ERROR_STATE_STD)
    return set_err(PDQ_ERR_INVALID, "unknown error estimator %d", cfg->error_estimator);
  if (cfg->error_derivative_idx < 0 || cfg->error_derivative_idx > cfg->num_derivatives)
    return set_err(PDQ_ERR_INVALID, "error derivative_idx %d outside [0, %d]", cfg->error_derivative_idx,
                   cfg->num_derivatives);
  if (cfg->prior != PDQ_PRIOR_IWP && cfg->prior != PDQ_PRIOR_EXPONENTIAL)
    return set_err(PDQ_ERR_INVALID, "unknown prior %d", cfg->prior);
  if (cfg->prior == PDQ_PRIOR_EXPONENTIAL) {
    if (cfg->factorisation != PDQ_DENSE)
      return set_err(PDQ_ERR_UNSUPPORTED,
                     "exponential priors are implemented for the dense factorisation (as in the reference, "
                     "ssm_impl_isotropic.py:484-520, ssm_impl_blockdiag.py:549-585)");
    if (!cfg->prior_drift) return set_err(PDQ_ERR_INVALID, "the exponential prior needs prior_drift");
    if ((cfg->num_derivatives + 1) * cfg->ode_dim > 16)
      return set_err(PDQ_ERR_UNSUPPORTED, "the exponential prior supports (nu + 1) d <= 16 (got %d)",
                     (cfg->num_derivatives + 1) * cfg->ode_dim);
  }
  if (cfg->jacobian < PDQ_JAC_EXACT || cfg->jacobian > PDQ_JAC_MONTE_CARLO_FWD)
    return set_err(PDQ_ERR_INVALID, "unknown jacobian handler %d", cfg->jacobian);
  if (cfg->jacobian != PDQ_JAC_EXACT) {
    if (cfg->jac_probes < 1 || cfg->jac_probes > 64)
      return set_err(PDQ_ERR_INVALID, "jac_probes must be in [1, 64] (got %d)", cfg->jac_probes);
    // re_linearize_before_error draws from the error estimator's own key
    // (solvers.py:946-950), which the probe sets do not model (the dense
    // factorisation materialises the exact Jacobian and draws none,
    // jacobians.py:221-228)
    if (cfg->constraint == PDQ_TS1 && cfg->factorisation != PDQ_DENSE && cfg->error_re_linearize)
      return set_err(PDQ_ERR_UNSUPPORTED,
                     "the Monte Carlo Jacobian handlers do not support re_linearize_before_error");
  }
  return device ? check_device() : PDQ_OK;
}

int lds_limit() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
  return v;
}

TuneKey tune_key(const pdq_config& c, int n_inst, int mode) {
  int bucket = 1;
  while (bucket < n_inst) bucket <<= 1;
  return TuneKey{mode, c.factorisation, c.num_derivatives, c.ode_dim, c.strategy, c.solver, c.constraint,
                 c.vector_field, c.prior, bucket};
}

// Heuristic lanes per instance: enough lanes that the batch fills the chip
// (>= 4 wavefronts per CU) without giving a group more lanes than its
// largest matrix dimension can use.
int heuristic_lanes(const pdq::Layout& L, int n_inst) {
  const int cus = pdq::num_cus();
  int G = L.nc >= 16 ? 64 : (L.nc >= 8 ? 32 : 16);
  while (G > 4 && (long)n_inst * G / 64 > 4L * cus * 4) G /= 2;
  while (G < 64 && (long)n_inst * G / 64 < 4L * cus) G *= 2;
  if (n_inst <= 64 / G) G = 64;
  const int limit = lds_limit();
  while (G < 64 && pdq::coop_lds_bytes(L, G) > (size_t)limit) G *= 2;
  return G;
}

// Lanes per instance: 1 = the lane-per-instance isotropic kernel (the
// default where it is instantiated, save_at and fixed-grid solves), 4..64 =
// the cooperative kernel with that group size.
int pick_lanes(const pdq_config& cfg, const pdq::Layout& L, int n_inst, int mode) {
  if (cfg.lanes_per_instance > 0) return cfg.lanes_per_instance;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(tune_key(cfg, n_inst, mode));
    if (it != g_tuned.end()) return it->second;
  }
  const bool lane_ok = cfg.factorisation == PDQ_ISOTROPIC && mode != pdq::MODE_LML && pdq::lane_supported(L.N, L.d);
  if (lane_ok) return 1;
  return heuristic_lanes(L, n_inst);
}

// Checks of a batched solve's configuration, inputs and outputs (before any
// device work).
int solve_validate(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs,
                   const double* times, int K, const pdq_solution* out) {
  int rc = check_common(cfg, n_inst, K);
  if (rc) return rc;
  if (!out) return set_err(PDQ_ERR_INVALID, "out is NULL");
  if (cfg->factorisation == PDQ_BLOCKDIAG) {
    if (cfg->ode_dim > pdq::kBdMaxDim)
      return set_err(PDQ_ERR_UNSUPPORTED, "blockdiag supports ode_dim <= %d (one lane per block; got %d)",
                     pdq::kBdMaxDim, cfg->ode_dim);
    if (cfg->num_derivatives > 5)
      return set_err(PDQ_ERR_UNSUPPORTED, "blockdiag supports num_derivatives in [1, 5] (got %d)", cfg->num_derivatives);
    if (cfg->init_std_rows)
      return set_err(PDQ_ERR_UNSUPPORTED, "init_std_rows (a std per row) is implemented for the dense and isotropic "
                                          "factorisations; blockdiag takes init_std");
  }
  if (n_inst == 0) return PDQ_OK;
  if (!tcoeffs || !times || (cfg->n_params > 0 && !params))
    return set_err(PDQ_ERR_INVALID, "tcoeffs, times and params must be device pointers");
  const pdq::Layout L = pdq::make_layout(*cfg, 0);
  if (L.nc > 32) return set_err(PDQ_ERR_UNSUPPORTED, "dense state dimension %d exceeds 32", L.nc);
  if (L.smoother && (!out->filt_mean || !out->filt_chol || !out->cond_A || !out->cond_mean || !out->cond_chol ||
                     !out->cond_tl || !out->cond_to))
    return set_err(PDQ_ERR_INVALID, "smoothers need filt_* and cond_* outputs (backward pass storage)");
  if (cfg->solver == PDQ_SOLVER_MLE && !out->num_steps)
    return set_err(PDQ_ERR_INVALID, "solver_mle needs the num_steps output");
  return PDQ_OK;
}

// Launch of a validated batched solve on the IWP matrices A, Q (device).
// queue: the cooperative kernel's work-queue counter; null allocates one for
// this call (stream ordered).
int solve_run(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs, const double* times,
              int K, const pdq_solution* out, const double* A, const double* Q, int* queue, hipStream_t st,
              int mode) {
  const pdq::Layout L = pdq::make_layout(*cfg, 0);
  if (cfg->factorisation == PDQ_BLOCKDIAG) {
    // one lane per block, one lane group per instance (pdq_bd.h)
    pdq::BdArgs ba{};
    ba.cfg = *cfg;
    ba.n_inst = n_inst;
    ba.K = K;
    ba.mode = mode == pdq::MODE_FIXED_GRID ? pdq::LMODE_FIXED_GRID : pdq::LMODE_SAVE_AT;
    ba.G = pdq::bd_group_lanes(cfg->ode_dim);
    ba.params = params;
    ba.tcoeffs = tcoeffs;
    ba.times = times;
    ba.Q = Q;
    ba.out = *out;
    return pdq::launch_bd(ba, st);
  }
  const int G = pick_lanes(*cfg, L, n_inst, mode);
  if (G == 1) {
    if (cfg->factorisation != PDQ_ISOTROPIC || !pdq::lane_supported(L.N, L.d))
      return set_err(PDQ_ERR_UNSUPPORTED,
                     "lanes_per_instance = 1 (lane-per-instance kernel) needs the isotropic factorisation with "
                     "nu = 3 or 4 and d = 2 (got nu = %d, d = %d)",
                     L.N - 1, L.d);
    pdq::LArgs la{};
    la.cfg = *cfg;
    la.n_inst = n_inst;
    la.K = K;
    la.mode = mode == pdq::MODE_FIXED_GRID ? pdq::LMODE_FIXED_GRID : pdq::LMODE_SAVE_AT;
    la.params = params;
    la.tcoeffs = tcoeffs;
    la.times = times;
    la.A = A;
    la.Q = Q;
    la.out = *out;
    return pdq::launch_lane(la, L.N, L.d, st);
  }
  int* q = queue;
  if (!q) HIP_TRY(hipMallocAsync((void**)&q, sizeof(int), st));
  pdq::KArgs a{};
  a.cfg = *cfg;
  a.lay = L;
  a.n_inst = n_inst;
  a.K = K;
  a.mode = mode;
  a.params = params;
  a.tcoeffs = tcoeffs;
  a.times = times;
  a.A = A;
  a.Q = Q;
  a.out = *out;
  a.queue = q;
  const int rc = pdq::launch_coop(a, G, st);
  if (!queue) (void)hipFreeAsync(q, st);
  return rc;
}

int solve_common(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs,
                 const double* times, int K, const pdq_solution* out, void* stream, int mode) {
  int rc = solve_validate(cfg, n_inst, params, tcoeffs, times, K, out);
  if (rc || n_inst == 0) return rc;
  DevMats m;
  rc = device_mats(*cfg, false, m);
  if (rc) return rc;
  return solve_run(cfg, n_inst, params, tcoeffs, times, K, out, m.A, m.Q, nullptr, (hipStream_t)stream, mode);
}

// ------------------------------------------------ no-allocation solve
// pdq_solve_adaptive_save_at_noalloc keeps what pdq_solve_adaptive_save_at
// allocates (the cooperative kernel's work-queue counter, and the IWP
// matrices that device_mats caches) in `scratch`, a fixed array in device
// memory, one per GPU for the life of the process:
//   bytes [0, 256)                 work-queue counter (int), zeroed by a kernel per call
//   bytes [256, 256 + 8 n^2)       A (n x n, n = the state dimension of one block)
//   bytes [256 + 8 n^2, ...)       Q
// Host code addresses it only through the pointer hipGetSymbolAddress returns
// for the current device (g_noalloc_ptr), never by the array's name.
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocMaxState = 32;  // n <= 32, the solve kernels' own limit
constexpr size_t kNoAllocQueueBytes = 256;
constexpr size_t kNoAllocScratchBytes =
    kNoAllocQueueBytes + 2 * sizeof(double) * kNoAllocMaxState * kNoAllocMaxState;  // 16 640 B
unsigned char* g_noalloc_ptr[kNoAllocMaxDevices];

// The matrices travel as kernel arguments, chunk by chunk, so that filling
// the workspace needs neither a host staging buffer nor a copy engine call.
constexpr int kFillChunk = 448;
struct FillChunk {
  double* dst;
  int n;
  double v[kFillChunk];
};

__global__ void scratch_fill_kernel(FillChunk c) {
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) c.dst[i] = c.v[i];
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// The current device's address of `scratch` (looked up once per device).
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return set_err(PDQ_ERR_UNSUPPORTED, "device %d: the no-allocation workspace serves devices 0..%d", dev,
                   kNoAllocMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(PDQ_ERR_HIP, "hipGetSymbolAddress returned no address for the workspace");
    g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return PDQ_OK;
}

int fill_device(double* dst, const double* src, int count, hipStream_t st) {
  for (int off = 0; off < count; off += kFillChunk) {
    FillChunk c;
    c.dst = dst + off;
    c.n = std::min(kFillChunk, count - off);
    std::memcpy(c.v, src + off, sizeof(double) * c.n);
    hipLaunchKernelGGL(scratch_fill_kernel, dim3(1), dim3(256), 0, st, c);
    HIP_TRY(hipGetLastError());
  }
  return PDQ_OK;
}

// One thread per instance.  Up to 512 coefficients ((num + 1) d) the
// normalised coefficients c_k = u^(k) / k! live in a private array; beyond,
// they are formed in place in the instance's output rows, with the same
// arithmetic, and scaled by k! in a last pass.
__global__ void taylor_kernel(int vf, int d, int np, int n_inst, const double* params, const double* y0, double t0,
                              int num, double* out) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= n_inst) return;
  const double* prm = params + (size_t)inst * np;
  double* o = out + (size_t)inst * (num + 1) * d;
  if ((num + 1) * d <= 8 * 64) {
    double c[8 * 64];
    for (int j = 0; j < d; ++j) c[j] = y0[(size_t)inst * d + j];
    for (int k = 0; k < num; ++k)
      for (int j = 0; j < d; ++j) c[(k + 1) * d + j] = pdq::taylor_rhs(vf, d, c, prm, t0, k, j) / (double)(k + 1);
    for (int k = 0; k <= num; ++k)
      for (int j = 0; j < d; ++j) o[(size_t)k * d + j] = pdq::factorial_d(k) * c[k * d + j];
    return;
  }
  for (int j = 0; j < d; ++j) o[j] = y0[(size_t)inst * d + j];
  for (int k = 0; k < num; ++k)
    for (int j = 0; j < d; ++j) o[(size_t)(k + 1) * d + j] = pdq::taylor_rhs(vf, d, (const double*)o, prm, t0, k, j) / (double)(k + 1);
  for (int k = 0; k <= num; ++k)
    for (int j = 0; j < d; ++j) o[(size_t)k * d + j] = pdq::factorial_d(k) * o[(size_t)k * d + j];
}

}  // namespace

extern "C" {

int32_t pdq_version(void) { return 1; }

int32_t pdq_system_matrices_1d_iwp(int32_t num_derivatives, double* A, double* Q) {
  if (num_derivatives < 0 || num_derivatives > 7)
    return set_err(PDQ_ERR_UNSUPPORTED, "num_derivatives must be in [0, 7] (got %d)", num_derivatives);
  if (!A || !Q) return set_err(PDQ_ERR_INVALID, "A and Q must be host pointers");
  std::vector<double> a, q;
  pdq::iwp_matrices(num_derivatives, a, q);
  std::memcpy(A, a.data(), a.size() * sizeof(double));
  std::memcpy(Q, q.data(), q.size() * sizeof(double));
  return PDQ_OK;
}

int32_t pdq_init(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_err(PDQ_ERR_NO_DEVICE, "no HIP device");
  unsigned char* ws = nullptr;
  return noalloc_scratch(&ws);  // the current device's no-allocation workspace
}

size_t pdq_last_error(char* buf, size_t size) {
  if (buf && size) {
    std::strncpy(buf, g_err, size - 1);
    buf[size - 1] = 0;
  }
  return std::strlen(g_err);
}

int32_t pdq_mean_size(const pdq_config* cfg) { return (cfg->num_derivatives + 1) * cfg->ode_dim; }

int32_t pdq_chol_size(const pdq_config* cfg) {
  const int N = cfg->num_derivatives + 1, d = cfg->ode_dim;
  if (cfg->factorisation == PDQ_DENSE) return N * d * N * d;
  if (cfg->factorisation == PDQ_ISOTROPIC) return N * N;
  return d * N * N;
}

int32_t pdq_solve_adaptive_save_at(const pdq_config* cfg, int32_t n_inst, const double* params,
                                   const double* tcoeffs, const double* save_at, int32_t K,
                                   const pdq_solution* out, void* stream) {
  return solve_common(cfg, n_inst, params, tcoeffs, save_at, K, out, stream, pdq::MODE_SAVE_AT);
}

int32_t pdq_solve_adaptive_save_at_noalloc(const pdq_config* cfg, int32_t n_inst, const double* params,
                                           const double* tcoeffs, const double* save_at, int32_t K,
                                           const pdq_solution* out, void* stream) {
  int rc = solve_validate(cfg, n_inst, params, tcoeffs, save_at, K, out);
  if (rc || n_inst == 0) return rc;
  const pdq::Layout L = pdq::make_layout(*cfg, 0);
  const int n = L.nc;
  if (n > kNoAllocMaxState)
    return set_err(PDQ_ERR_UNSUPPORTED,
                   "the no-allocation solve holds state dimensions up to %d in its %zu-byte workspace (got %d)",
                   kNoAllocMaxState, kNoAllocScratchBytes, n);
  unsigned char* ws = nullptr;
  rc = noalloc_scratch(&ws);
  if (rc) return rc;
  double hA[kNoAllocMaxState * kNoAllocMaxState], hQ[kNoAllocMaxState * kNoAllocMaxState];
  if (cfg->factorisation == PDQ_DENSE)
    pdq::iwp_dense_into(cfg->num_derivatives, cfg->ode_dim, cfg->base_output_scale, hA, hQ);
  else
    pdq::iwp_matrices_into(cfg->num_derivatives, hA, hQ);
  double* A = reinterpret_cast<double*>(ws + kNoAllocQueueBytes);
  double* Q = A + (size_t)n * n;
  hipStream_t st = (hipStream_t)stream;
  rc = fill_device(A, hA, n * n, st);
  if (!rc) rc = fill_device(Q, hQ, n * n, st);
  if (rc) return rc;
  // the queue counter is zeroed by the cooperative launcher's own kernel
  return solve_run(cfg, n_inst, params, tcoeffs, save_at, K, out, A, Q, reinterpret_cast<int*>(ws), st,
                   pdq::MODE_SAVE_AT);
}

int32_t pdq_solve_fixed_grid(const pdq_config* cfg, int32_t n_inst, const double* params, const double* tcoeffs,
                             const double* grid, int32_t K, const pdq_solution* out, void* stream) {
  return solve_common(cfg, n_inst, params, tcoeffs, grid, K, out, stream, pdq::MODE_FIXED_GRID);
}

int32_t pdq_lml_value_and_grad(const pdq_config* cfg, int32_t n_inst, const double* params, const double* y0,
                               const double* grid, int32_t K_grid, int32_t every, const double* data,
                               const double* stds, int32_t n_std, int32_t average_pdfs, double* value, double* grad,
                               void* stream) {
  // the device is checked after the LML's own configuration checks
  int rc = check_common(cfg, n_inst, K_grid, false);
  if (rc) return rc;
  if (cfg->factorisation == PDQ_BLOCKDIAG)
    return set_err(PDQ_ERR_UNSUPPORTED, "the LML path supports the dense and isotropic factorisations");
  if (cfg->strategy != PDQ_FIXEDPOINT)
    return set_err(PDQ_ERR_UNSUPPORTED, "the LML needs strategy_smoother_fixedpoint (a Markov sequence)");
  if (cfg->constraint != PDQ_TS0 && grad)
    return set_err(PDQ_ERR_UNSUPPORTED, "gradients need ts0 (the covariance recursion must not depend on them)");
  // exponential priors (dense only, check_common) run the cooperative LML: their
  // transition depends on the step and the drift, not on the parameters, so the
  // forward-mode tangents stay exact
  // solver_mle: the value runs (the posterior rescaled by the calibrated output
  // scale); its gradient would carry that scale's dependence on the parameters
  // through every factor of the reverse scan, which the tangents do not
  if (cfg->solver == PDQ_SOLVER_MLE && grad)
    return set_err(PDQ_ERR_UNSUPPORTED,
                   "LML gradients with solver_mle are not supported (the value runs: pass grad = NULL)");
  if (every < 1 || (K_grid - 1) % every != 0)
    return set_err(PDQ_ERR_INVALID, "every (%d) must divide K_grid - 1 (%d)", every, K_grid - 1);
  const int n_tan = grad ? cfg->n_params + cfg->ode_dim : 0;
  if (n_tan > pdq::kMaxTangents) return set_err(PDQ_ERR_UNSUPPORTED, "at most %d gradient directions", pdq::kMaxTangents);
  if (cfg->factorisation == PDQ_DENSE && cfg->ode_dim > 1 && n_std != 1 && n_std != cfg->ode_dim)
    return set_err(PDQ_ERR_INVALID, "n_std must be 1 or ode_dim");
  if (cfg->factorisation == PDQ_ISOTROPIC && n_std != 1)
    return set_err(PDQ_ERR_INVALID, "isotropic observation noise is one scalar per time point (n_std = 1)");
  if ((cfg->num_derivatives + 1) * cfg->ode_dim > 64) return set_err(PDQ_ERR_UNSUPPORTED, "Taylor init too large");
  const int N = cfg->num_derivatives + 1, d = cfg->ode_dim;
  if (cfg->lml_tcoeff_index < 0 || cfg->lml_tcoeff_index > cfg->num_derivatives)
    return set_err(PDQ_ERR_INVALID, "lml_tcoeff_index (%d) must lie in [0, num_derivatives]", cfg->lml_tcoeff_index);
  // the lane LML's QR is specialised to observations of coefficient 0 and its
  // forward pass to ts0; the other cases run the cooperative LML
  const bool lane_ok = cfg->factorisation == PDQ_ISOTROPIC && cfg->lml_tcoeff_index == 0 &&
                       cfg->constraint == PDQ_TS0 && pdq::lane_lml_supported(N, d, cfg->solver);
  if (cfg->lanes_per_instance == 1 && !lane_ok)
    return set_err(PDQ_ERR_UNSUPPORTED,
                   "lanes_per_instance = 1 (lane-per-instance LML) needs the isotropic factorisation with nu = 3 or "
                   "4, d = 2, ts0, solver or solver_dynamic and lml_tcoeff_index = 0 (got nu = %d, d = %d)",
                   N - 1, d);
  if (cfg->lml_solve != PDQ_LML_SOLVE_LSTSQ && cfg->lml_solve != PDQ_LML_SOLVE_TRIU)
    return set_err(PDQ_ERR_INVALID, "unknown lml_solve %d", cfg->lml_solve);
  if (cfg->lml_gradient != PDQ_GRAD_FORWARD && cfg->lml_gradient != PDQ_GRAD_REVERSE)
    return set_err(PDQ_ERR_INVALID, "unknown lml_gradient %d", cfg->lml_gradient);
  const bool reverse = grad && cfg->lml_gradient == PDQ_GRAD_REVERSE;
  if (reverse && !lane_ok)
    return set_err(PDQ_ERR_UNSUPPORTED,
                   "the reverse adjoint (lml_gradient = PDQ_GRAD_REVERSE) covers the isotropic factorisation with "
                   "nu = 3 or 4, d = 2, ts0, solver or solver_dynamic and lml_tcoeff_index = 0 (got nu = %d, d = %d)",
                   N - 1, d);
  if (reverse && cfg->lanes_per_instance > 1)
    return set_err(PDQ_ERR_UNSUPPORTED, "the reverse adjoint runs one lane per instance (lanes_per_instance 0 or 1)");
  rc = check_device();
  if (rc) return rc;
  if (n_inst == 0) return PDQ_OK;
  if (!params || !y0 || !grid || !data || !stds) return set_err(PDQ_ERR_INVALID, "null input pointer");
  hipStream_t st = (hipStream_t)stream;
  const int Kobs = (K_grid - 1) / every + 1;
  if (reverse) {
    DevMats m;
    rc = device_mats(*cfg, false, m);
    if (rc) return rc;
    CallBuffers bufs(st);
    double* tinit = nullptr;
    double* adj_buf = nullptr;  // conditionals, node data, checkpoints and one segment per lane
    HIP_TRY(bufs.alloc(&tinit, sizeof(double) * (size_t)n_inst * N * d * (1 + n_tan)));
    HIP_TRY(bufs.alloc(&adj_buf, sizeof(double) * pdq::lane_adj_scratch_doubles(N, d, K_grid, every, n_inst)));
    rc = pdq::launch_taylor_tangents(cfg->vector_field, d, cfg->n_params, n_inst, cfg->num_derivatives, n_tan, params,
                                     y0, grid, tinit, st);
    if (rc) return rc;
    pdq::LmlArgs la{};
    la.cfg = *cfg;
    la.n_inst = n_inst;
    la.K = K_grid;
    la.every = every;
    la.n_tan = n_tan;
    la.average_pdfs = average_pdfs;
    la.G = 1;
    la.params = params;
    la.tcoeffs = tinit;
    la.times = grid;
    la.data = data;
    la.stds = stds;
    la.Q = m.Q;
    la.value = value;
    la.grad = grad;
    la.scratch = adj_buf;
    return pdq::launch_lane_adj(la, N, d, st);
  }
  if (lane_ok && cfg->lanes_per_instance <= 1) {
    // lane-per-instance LML: G lanes per instance, `tpl` tangent columns each
    const int tpl = pdq::lane_lml_tangents_per_lane(n_tan);
    const int G = tpl ? n_tan / tpl : 1;
    const size_t n_lanes = (size_t)n_inst * G;
    const size_t csz = pdq::lane_lml_cond_size(N, d, tpl);
    DevMats m;
    rc = device_mats(*cfg, false, m);
    if (rc) return rc;
    CallBuffers bufs(st);
    double* tinit = nullptr;
    double* cond_buf = nullptr;  // backward conditionals between observation nodes
    HIP_TRY(bufs.alloc(&tinit, sizeof(double) * (size_t)n_inst * N * d * (1 + n_tan)));
    HIP_TRY(bufs.alloc(&cond_buf, sizeof(double) * csz * (size_t)std::max(1, Kobs - 1) * n_lanes));
    rc = pdq::launch_taylor_tangents(cfg->vector_field, d, cfg->n_params, n_inst, cfg->num_derivatives, n_tan, params,
                                     y0, grid, tinit, st);
    if (rc) return rc;
    pdq::LmlArgs la{};
    la.cfg = *cfg;
    la.n_inst = n_inst;
    la.K = K_grid;
    la.every = every;
    la.n_tan = n_tan;
    la.average_pdfs = average_pdfs;
    la.G = G;
    la.params = params;
    la.tcoeffs = tinit;
    la.times = grid;
    la.data = data;
    la.stds = stds;
    la.Q = m.Q;
    la.value = value;
    la.grad = grad;
    la.scratch = cond_buf;
    return pdq::launch_lane_lml(la, N, d, tpl, st);
  }
  pdq::Layout L = pdq::make_layout(*cfg, n_tan);
  if (L.nc > 32) return set_err(PDQ_ERR_UNSUPPORTED, "dense state dimension %d exceeds 32", L.nc);
  const size_t csz = 2 * (size_t)L.nc * L.nc + (size_t)L.nc * L.mc + 2 * (size_t)L.nc;
  DevMats m;
  rc = device_mats(*cfg, false, m);
  if (rc) return rc;
  CallBuffers bufs(st);
  double* cond_buf = nullptr;  // backward conditionals between observation nodes
  int* queue = nullptr;
  double* tinit = nullptr;  // prior means with tangents: Taylor expansion of y0 (jetexpand, num = nu)
  HIP_TRY(bufs.alloc(&queue, sizeof(int)));
  HIP_TRY(bufs.alloc(&cond_buf, sizeof(double) * csz * (size_t)std::max(1, Kobs - 1) * n_inst));
  HIP_TRY(bufs.alloc(&tinit, sizeof(double) * (size_t)n_inst * N * d * (1 + n_tan)));
  rc = pdq::launch_taylor_tangents(cfg->vector_field, d, cfg->n_params, n_inst, cfg->num_derivatives, n_tan, params, y0,
                                   grid, tinit, st);
  if (rc) return rc;
  pdq::KArgs a{};
  a.cfg = *cfg;
  a.lay = L;
  a.n_inst = n_inst;
  a.K = K_grid;
  a.mode = pdq::MODE_LML;
  a.every = every;
  a.n_tan = n_tan;
  a.average_pdfs = average_pdfs;
  a.n_std = n_std;
  a.params = params;
  a.tcoeffs = tinit;
  a.times = grid;
  a.data = data;
  a.stds = stds;
  a.A = m.A;
  a.Q = m.Q;
  a.lml_value = value;
  a.lml_grad = grad;
  a.scratch = cond_buf;
  a.queue = queue;
  const int G = pick_lanes(*cfg, L, n_inst, pdq::MODE_LML);
  return pdq::launch_coop(a, G, st);
}

int32_t pdq_offgrid_marginals(const pdq_config* cfg, int32_t n_inst, int32_t K, const double* times,
                              const pdq_solution* solution, int32_t n_query, const double* t_query,
                              double* mean_out, double* chol_out, void* stream) {
  int rc = check_common(cfg, n_inst, K);
  if (rc) return rc;
  if (!solution) return set_err(PDQ_ERR_INVALID, "solution is NULL");
  if (n_query < 0) return set_err(PDQ_ERR_INVALID, "n_query < 0");
  if (cfg->strategy == PDQ_FIXEDPOINT)
    return set_err(PDQ_ERR_UNSUPPORTED,
                   "offgrid marginals need a filter or a fixed-interval smoother (the fixed-point smoother is not "
                   "suitable, solvers.py:163-164)");
  const int N = cfg->num_derivatives + 1, d = cfg->ode_dim;
  if (cfg->factorisation == PDQ_BLOCKDIAG)
    return set_err(PDQ_ERR_UNSUPPORTED, "offgrid marginals are implemented for the dense and isotropic factorisations");
  const bool lane = cfg->factorisation == PDQ_ISOTROPIC && pdq::lane_supported(N, d) && cfg->lanes_per_instance <= 1;
  pdq::Layout L = pdq::make_layout(*cfg, 0);
  if (!lane && L.nc > 32) return set_err(PDQ_ERR_UNSUPPORTED, "dense state dimension %d exceeds 32", L.nc);
  if (n_inst == 0 || n_query == 0) return PDQ_OK;
  if (!times || !t_query || !mean_out || !chol_out || !solution->mean || !solution->chol || !solution->output_scale)
    return set_err(PDQ_ERR_INVALID, "times, t_query, the outputs and the solution's mean, chol and output_scale "
                                    "must be device pointers");
  if (cfg->strategy == PDQ_FIXEDINTERVAL && (!solution->filt_mean || !solution->filt_chol))
    return set_err(PDQ_ERR_INVALID, "the fixed-interval smoother needs the solution's filtering marginals");
  DevMats m;
  rc = device_mats(*cfg, false, m);
  if (rc) return rc;
  if (!lane) {
    // the cooperative kernel's primitives, one G-lane group per query
    pdq::OffgridCoopArgs o{};
    o.k.cfg = *cfg;
    o.k.lay = L;
    o.k.n_inst = n_inst;
    o.k.K = K;
    o.k.times = times;
    o.k.A = m.A;
    o.k.Q = m.Q;
    o.n_query = n_query;
    o.t_query = t_query;
    o.mean = solution->mean;
    o.chol = solution->chol;
    o.filt_mean = solution->filt_mean;
    o.filt_chol = solution->filt_chol;
    o.output_scale = solution->output_scale;
    o.mean_out = mean_out;
    o.chol_out = chol_out;
    int G = cfg->lanes_per_instance > 1 ? cfg->lanes_per_instance : (L.nc <= 8 ? 8 : (L.nc <= 16 ? 16 : 32));
    const int limit = lds_limit();
    while (G < 64 && pdq::coop_lds_bytes(L, G) > (size_t)limit) G *= 2;
    return pdq::launch_offgrid_coop(o, G, (hipStream_t)stream);
  }
  pdq::OffgridArgs a{};
  a.cfg = *cfg;
  a.n_inst = n_inst;
  a.K = K;
  a.n_query = n_query;
  a.times = times;
  a.t_query = t_query;
  a.mean = solution->mean;
  a.chol = solution->chol;
  a.filt_mean = solution->filt_mean;
  a.filt_chol = solution->filt_chol;
  a.output_scale = solution->output_scale;
  a.Q = m.Q;
  a.mean_out = mean_out;
  a.chol_out = chol_out;
  return pdq::launch_offgrid(a, N, d, (hipStream_t)stream);
}

int32_t pdq_sample(const pdq_config* cfg, int32_t n_inst, int32_t K, const pdq_solution* solution,
                   int32_t n_samples, const double* base, double* samples, void* stream) {
  int rc = check_common(cfg, n_inst, K);
  if (rc) return rc;
  if (!solution) return set_err(PDQ_ERR_INVALID, "solution is NULL");
  if (n_samples < 0) return set_err(PDQ_ERR_INVALID, "n_samples < 0");
  if (cfg->strategy == PDQ_FILTER)
    return set_err(PDQ_ERR_UNSUPPORTED, "sampling needs a smoother's posterior (a filter's solution is no Markov sequence)");
  const int N = cfg->num_derivatives + 1, d = cfg->ode_dim;
  if (cfg->factorisation == PDQ_BLOCKDIAG)
    return set_err(PDQ_ERR_UNSUPPORTED, "sampling is implemented for the dense and isotropic factorisations");
  const bool lane = cfg->factorisation == PDQ_ISOTROPIC && pdq::lane_supported(N, d) && cfg->lanes_per_instance <= 1;
  const pdq::Layout L = pdq::make_layout(*cfg, 0);
  if (!lane && L.nc * L.mc > pdq::kSampleGenericMax)
    return set_err(PDQ_ERR_UNSUPPORTED, "sampling supports (nu + 1) d <= %d (got %d)", pdq::kSampleGenericMax,
                   L.nc * L.mc);
  if (n_inst == 0 || n_samples == 0) return PDQ_OK;
  const pdq_solution& s = *solution;
  if (!base || !samples || !s.term_mean || !s.term_chol || !s.cond_A || !s.cond_mean || !s.cond_chol || !s.cond_tl ||
      !s.cond_to)
    return set_err(PDQ_ERR_INVALID, "base, samples and the solution's term_* and cond_* must be device pointers");
  pdq::SampleArgs a{};
  a.cfg = *cfg;
  a.n_inst = n_inst;
  a.K = K;
  a.n_samples = n_samples;
  a.term_mean = s.term_mean;
  a.term_chol = s.term_chol;
  a.cond_A = s.cond_A;
  a.cond_mean = s.cond_mean;
  a.cond_chol = s.cond_chol;
  a.cond_tl = s.cond_tl;
  a.cond_to = s.cond_to;
  a.base = base;
  a.samples = samples;
  if (!lane)
    return pdq::launch_sample_generic(a, L.nc, L.mc, cfg->factorisation == PDQ_DENSE ? 1 : 0, (hipStream_t)stream);
  return pdq::launch_sample(a, N, d, (hipStream_t)stream);
}

int32_t pdq_lml_terminal_values(const pdq_config* cfg, int32_t n_inst, const double* mean, const double* chol,
                                const double* u, const double* std, int32_t tcoeff_index, double* value,
                                void* stream) {
  if (!cfg) return set_err(PDQ_ERR_INVALID, "cfg is NULL");
  if (n_inst < 0) return set_err(PDQ_ERR_INVALID, "n_inst < 0");
  if (cfg->factorisation == PDQ_BLOCKDIAG)
    return set_err(PDQ_ERR_UNSUPPORTED, "loss_lml_terminal_values is implemented for the isotropic and dense factorisations");
  if (cfg->num_derivatives < 0 || cfg->num_derivatives > 7 || cfg->ode_dim < 1)
    return set_err(PDQ_ERR_INVALID, "invalid num_derivatives or ode_dim");
  if (tcoeff_index < 0 || tcoeff_index > cfg->num_derivatives)
    return set_err(PDQ_ERR_INVALID, "tcoeff_index %d outside [0, %d]", tcoeff_index, cfg->num_derivatives);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(PDQ_ERR_NO_DEVICE, "no HIP device");
  if (n_inst == 0) return PDQ_OK;
  if (!mean || !chol || !u || !std || !value) return set_err(PDQ_ERR_INVALID, "null device pointer");
  return pdq::launch_terminal_lml(*cfg, n_inst, tcoeff_index, mean, chol, u, std, value, (hipStream_t)stream);
}

int32_t pdq_taylor_coefficients(int32_t vector_field, int32_t d, int32_t n_params, int32_t n_inst,
                                const double* params, const double* y0, double t0, int32_t num, double* out,
                                void* stream) {
  if (num < 0 || d < 1 || n_inst < 0) return set_err(PDQ_ERR_INVALID, "invalid sizes");
  if (num > 7) return set_err(PDQ_ERR_UNSUPPORTED, "num must not exceed 7 (got %d)", num);
  if (d > (1 << 20)) return set_err(PDQ_ERR_UNSUPPORTED, "d must not exceed 2^20 (got %d)", d);
  const int rc = check_vector_field(vector_field, d, n_params);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(PDQ_ERR_NO_DEVICE, "no HIP device");
  if (n_inst == 0) return PDQ_OK;
  if (!y0 || !out || (n_params > 0 && !params)) return set_err(PDQ_ERR_INVALID, "null device pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(taylor_kernel, dim3((n_inst + 63) / 64), dim3(64), 0, st, vector_field, d, n_params, n_inst,
                     params, y0, t0, num, out);
  HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

int32_t pdq_tuned_lanes(const pdq_config* cfg, int32_t n_inst) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tuned.find(tune_key(*cfg, n_inst, pdq::MODE_SAVE_AT));
  return it == g_tuned.end() ? 0 : it->second;
}

// Records the lanes-per-instance choice of an external timing sweep.
int32_t pdq_set_tuned_lanes(const pdq_config* cfg, int32_t n_inst, int32_t mode, int32_t lanes) {
  if (lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
    return set_err(PDQ_ERR_INVALID, "lanes must be 1 (lane kernel) or 4, 8, 16, 32, 64 (cooperative kernel)");
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuned[tune_key(*cfg, n_inst, mode)] = lanes;
  return PDQ_OK;
}

}  // extern "C"
