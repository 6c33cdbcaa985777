// CPU build of the kernels' per-instance code (test infrastructure).
// Compiled with g++ -DPDQ_HOST_EMU from the product's device sources; runs
// every instance with one lane (G = 1) so that the kernel logic can be
// compared with the oracle on the CPU.
#include <barrier>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../probdiffeq_amd/csrc/pdq_bd.h"
#include "../../probdiffeq_amd/csrc/pdq_coop.hip"
#include "../../probdiffeq_amd/csrc/pdq_iwp.h"
#include "../../probdiffeq_amd/csrc/pdq_lane.h"
#include "../../probdiffeq_amd/csrc/pdq_lane_adjoint.h"
#include "../../probdiffeq_amd/csrc/pdq_lane_lml.h"
#include "../../probdiffeq_amd/csrc/pdq_lane_offgrid.h"
#include "../../probdiffeq_amd/csrc/pdq_lane_sample.h"
#include "../../probdiffeq_amd/csrc/pdq_sample_generic.h"

namespace pdq {
int host_error(int code, const char*, ...) { return code; }
int num_cus() { return 1; }
}  // namespace pdq

// The block-diagonal kernel's lane group on the CPU: one thread per lane, a
// real barrier, and lane exchanges through a shared array with HIP's
// __shfl_xor / __shfl semantics (a source lane outside the caller's width
// segment reads the caller's own value).
namespace {
std::barrier<>* g_group_barrier = nullptr;
double g_group_xchg[pdq::kBdMaxDim];
thread_local int t_group_lane = 0;
}  // namespace

void pdq_emu_group_barrier() { g_group_barrier->arrive_and_wait(); }

double pdq_emu_shfl_xor(double v, int mask, int width) {
  const int self = t_group_lane;
  g_group_xchg[self] = v;
  g_group_barrier->arrive_and_wait();
  int idx = self ^ mask;
  if (idx >= ((self + width) & ~(width - 1))) idx = self;
  const double r = g_group_xchg[idx];
  g_group_barrier->arrive_and_wait();
  return r;
}

double pdq_emu_shfl(double v, int src, int width) {
  const int self = t_group_lane;
  g_group_xchg[self] = v;
  g_group_barrier->arrive_and_wait();
  const double r = g_group_xchg[(self & ~(width - 1)) + src];
  g_group_barrier->arrive_and_wait();
  return r;
}

namespace {
template <int N, int S>
void bd_emu_instance(const pdq::BdArgs& a, int inst) {
  const int G = a.G;
  std::vector<double> y(2 * G, 0.0), red(2 * pdq::kBdRedStride, 0.0);
  std::barrier<> bar(G);
  g_group_barrier = &bar;
  std::vector<std::thread> th;
  th.reserve(G);
  for (int l = 0; l < G; ++l)
    th.emplace_back([&, l] {
      t_group_lane = l;
      pdq::BdGroup g;
      g.G = G;
      g.lane = l;
      g.multi = G > 64;
      g.y = y.data();
      g.red = red.data();
      pdq::bd_solve_instance<N, S>(a, g, inst);
    });
  for (auto& t : th) t.join();
  g_group_barrier = nullptr;
}

template <int N>
void bd_emu_n(const pdq::BdArgs& a, int inst) {
  switch (a.cfg.strategy) {
    case PDQ_FILTER: return bd_emu_instance<N, PDQ_FILTER>(a, inst);
    case PDQ_FIXEDPOINT: return bd_emu_instance<N, PDQ_FIXEDPOINT>(a, inst);
    default: return bd_emu_instance<N, PDQ_FIXEDINTERVAL>(a, inst);
  }
}
}  // namespace

extern "C" {

// Entries where the lane kernel's compile-time IWP transition (pdq_lane.h:
// lane_iwp_a) differs from the host's iwp_matrices, for nu = 1..7.
int pdqemu_iwp_mismatches(void) {
  int bad = 0;
  auto check = [&](auto n_tag) {
    constexpr int N = decltype(n_tag)::value;
    std::vector<double> A, Q;
    pdq::iwp_matrices(N - 1, A, Q);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) bad += pdq::lane_iwp_a<N>(i, j) != A[i * N + j];
    for (int i = 0; i < N; ++i)
      for (int j = i + 1; j < N; ++j) bad += Q[i * N + j] != 0.0;  // Q lower triangular
  };
  check(std::integral_constant<int, 2>{});
  check(std::integral_constant<int, 3>{});
  check(std::integral_constant<int, 4>{});
  check(std::integral_constant<int, 5>{});
  check(std::integral_constant<int, 6>{});
  check(std::integral_constant<int, 7>{});
  check(std::integral_constant<int, 8>{});
  return bad;
}

// Entries (bitwise) where lane_revert_left followed by lane_revert_right
// differs from lane_revert_prior (pdq_lane.h), over `trials` random reverts
// of the IWP prior with nu + 1 = N = 2..6 and D = 2 (the isotropic solve) or
// D = 4 (the LML's primal and tangent columns): prediction mean and factor
// and every entry of the conditional.  Trial 0 reverts a zero factor (no
// reflections at all), trial 1 a factor with a zero column.
int pdqemu_revert_halves_mismatch(int N_, int D_, int trials, unsigned seed, int rescale) {
  int bad = 0;
  unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
  auto unif = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * 0x1.0p-53;
  };
  auto check = [&](auto n_tag, auto d_tag) {
    constexpr int N = decltype(n_tag)::value, D = decltype(d_tag)::value;
    using Cond = pdq::LCond<N, D>;
    std::vector<double> A, Qv;
    pdq::iwp_matrices(N - 1, A, Qv);
    for (int t = 0; t < trials; ++t) {
      const double h = 1e-3 + unif() * 0.5, s = 0.1 + unif() * 10.0;
      double P[N], PI[N], Xm[N * D], XL[N * N];
      pdq::lane_precon<N>(h, P, PI);
      for (int i = 0; i < N * D; ++i) Xm[i] = 4.0 * unif() - 2.0;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          XL[i * N + j] = (j <= i && t != 0 && !(t == 1 && j == 1)) ? (2.0 * unif() - 1.0) * 1e-2 : 0.0;
      double Ym0[N * D], YL0[N * N], Ym1[N * D], YL1[N * N];
      Cond c0;
      std::memset(&c0, 0, sizeof c0);
      pdq::lane_revert_prior<N, D>(Qv.data(), P, PI, s, Xm, XL, Ym0, YL0, c0, rescale != 0);
      // as in the kernel: the stage in the conditional's own storage (lane stride) and a tail
      std::vector<double> cbuf(pdq::CondLayout<N, D>::size * pdq::kLaneBlock, 0.0);
      std::vector<double> tail(pdq::RevertStage<N, D>::tail + 1, 0.0);
      const pdq::StageRef<N, D> st{cbuf.data(), tail.data(), 1};
      const pdq::CondRef<N, D> c1{cbuf.data()};
      pdq::lane_revert_left<N, D>(Qv.data(), P, PI, s, Xm, XL, Ym1, YL1, st);
      pdq::lane_revert_right<N, D>(Xm, st, c1, rescale != 0);
      bad += std::memcmp(Ym0, Ym1, sizeof Ym0) != 0;
      bad += std::memcmp(YL0, YL1, sizeof YL0) != 0;
      for (int e = 0; e < pdq::CondLayout<N, D>::size; ++e) {
        if (e >= pdq::CondLayout<N, D>::L && e < pdq::CondLayout<N, D>::tl) continue;  // L: below
        bad += std::memcmp(&c0.v[e], &c1.at(e), sizeof(double)) != 0;
      }
      for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) {
          const double x0 = c0.L(i, j), x1 = c1.L(i, j);
          bad += std::memcmp(&x0, &x1, sizeof(double)) != 0;
        }
    }
  };
  auto with_d = [&](auto n_tag) {
    if (D_ == 2) check(n_tag, std::integral_constant<int, 2>{});
    else if (D_ == 4) check(n_tag, std::integral_constant<int, 4>{});
    else bad = -1;
  };
  switch (N_) {
    case 2: with_d(std::integral_constant<int, 2>{}); break;
    case 3: with_d(std::integral_constant<int, 3>{}); break;
    case 4: with_d(std::integral_constant<int, 4>{}); break;
    case 5: with_d(std::integral_constant<int, 5>{}); break;
    case 6: with_d(std::integral_constant<int, 6>{}); break;
    default: return -1;
  }
  return bad;
}

// The lane kernel's shared-divisor quotient (pdq_lane.h: lane_div).
double pdqemu_lane_div(double a, double b) { return pdq::lane_div(a, b, 1.0 / b); }

// The lane kernel's dt^k (pdq_lane.h: lane_ipow) for k in [-7, 7].
double pdqemu_ipow(int k, double x) {
  switch (k) {
    case -7: return pdq::lane_ipow<-7>(x);
    case -6: return pdq::lane_ipow<-6>(x);
    case -5: return pdq::lane_ipow<-5>(x);
    case -4: return pdq::lane_ipow<-4>(x);
    case -3: return pdq::lane_ipow<-3>(x);
    case -2: return pdq::lane_ipow<-2>(x);
    case -1: return pdq::lane_ipow<-1>(x);
    case 0: return pdq::lane_ipow<0>(x);
    case 1: return pdq::lane_ipow<1>(x);
    case 2: return pdq::lane_ipow<2>(x);
    case 3: return pdq::lane_ipow<3>(x);
    case 4: return pdq::lane_ipow<4>(x);
    case 5: return pdq::lane_ipow<5>(x);
    case 6: return pdq::lane_ipow<6>(x);
    case 7: return pdq::lane_ipow<7>(x);
    default: return 0.0;
  }
}

// The lane-per-instance isotropic kernel's per-instance program (pdq_lane.h)
// on the CPU, one instance after another.  Returns 2 for a shape the device
// library does not instantiate.
int pdqemu_solve_lane(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs,
                      const double* times, int K, const pdq_solution* out, int mode) {
  const int N = cfg->num_derivatives + 1, D = cfg->ode_dim;
  if (cfg->factorisation != PDQ_ISOTROPIC) return 2;
  std::vector<double> A, Q;
  pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  pdq::LArgs a{};
  a.cfg = *cfg;
  a.n_inst = n_inst;
  a.K = K;
  a.mode = mode;
  a.params = params;
  a.tcoeffs = tcoeffs;
  a.times = times;
  a.A = A.data();
  a.Q = Q.data();
  a.out = *out;
  // the lane's slots (its LDS column on the device), element stride 64
  std::vector<double> slots(64 * 3 * 128);
  for (int i = 0; i < n_inst; ++i) {
    if (N == 4 && D == 2) pdq::lane_solve_instance<4, 2>(a, i, slots.data());
    else if (N == 5 && D == 2) pdq::lane_solve_instance<5, 2>(a, i, slots.data());
    else return 2;
  }
  return 0;
}

// mode: 0 save_at, 1 fixed grid.  Output pointers as in pdq_solution (host).
int pdqemu_solve(const pdq_config* cfg, int n_inst, const double* params, const double* tcoeffs, const double* times,
                 int K, const pdq_solution* out, int mode) {
  int queue = 0;
  if (cfg->factorisation == PDQ_BLOCKDIAG) {
    // the block-diagonal kernel's code, one instance at a time, each as a
    // group of G threads (one per block)
    std::vector<double> A1, Q1;
    pdq::iwp_matrices(cfg->num_derivatives, A1, Q1);
    pdq::BdArgs a{};
    a.cfg = *cfg;
    a.n_inst = n_inst;
    a.K = K;
    a.mode = mode == 1 ? pdq::LMODE_FIXED_GRID : pdq::LMODE_SAVE_AT;
    a.G = pdq::bd_group_lanes(cfg->ode_dim);
    a.params = params;
    a.tcoeffs = tcoeffs;
    a.times = times;
    a.Q = Q1.data();
    a.out = *out;
    if (cfg->ode_dim > pdq::kBdMaxDim) return 2;
    for (int i = 0; i < n_inst; ++i) {
      switch (cfg->num_derivatives + 1) {
        case 2: bd_emu_n<2>(a, i); break;
        case 3: bd_emu_n<3>(a, i); break;
        case 4: bd_emu_n<4>(a, i); break;
        case 5: bd_emu_n<5>(a, i); break;
        case 6: bd_emu_n<6>(a, i); break;
        default: return 2;
      }
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


// LML value and forward-mode gradient (MODE_LML) on the CPU.
}  // extern "C"

// Initial Taylor coefficients with tangents, as taylor_tangent_kernel computes them.
static std::vector<double> emu_taylor_tangents(const pdq_config* cfg, int n_inst, const double* params, const double* y0,
                                               const double* grid, int n_tan) {
  const int N = cfg->num_derivatives + 1, d = cfg->ode_dim, np_ = cfg->n_params;
  std::vector<double> tinit((size_t)n_inst * N * d * (1 + n_tan));
  for (int inst = 0; inst < n_inst; ++inst) {
    pdq::Dual prmD[pdq::kMaxParams], cc[64];
    for (int q = 0; q < np_; ++q) {
      prmD[q] = pdq::dconst(params[(size_t)inst * np_ + q]);
      if (q < n_tan) prmD[q].g[q] = 1.0;
    }
    for (int k = 0; k < d; ++k) {
      cc[k] = pdq::dconst(y0[(size_t)inst * d + k]);
      if (np_ + k < n_tan) cc[k].g[np_ + k] = 1.0;
    }
    for (int k = 0; k + 1 < N; ++k)
      for (int j = 0; j < d; ++j)
        cc[(k + 1) * d + j] = pdq::taylor_rhs(cfg->vector_field, d, cc, prmD, grid[0], k, j) / (double)(k + 1);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < d; ++k) {
        const pdq::Dual v = pdq::factorial_d(i) * cc[i * d + k];
        for (int q = 0; q <= n_tan; ++q)
          tinit[(((size_t)inst * N + i) * d + k) * (1 + n_tan) + q] = q == 0 ? v.v : v.g[q - 1];
      }
  }
  return tinit;
}

extern "C" {

int pdqemu_lml(const pdq_config* cfg, int n_inst, const double* params, const double* y0, const double* grid, int K,
               int every, const double* data, const double* stds, int n_std, int average, double* value,
               double* grad) {
  const int n_tan = grad ? cfg->n_params + cfg->ode_dim : 0;
  const int d = cfg->ode_dim;
  pdq::Layout L = pdq::make_layout(*cfg, n_tan);
  std::vector<double> A, Q;
  if (cfg->factorisation == PDQ_DENSE) pdq::iwp_dense(cfg->num_derivatives, d, cfg->base_output_scale, A, Q);
  else pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  std::vector<double> tinit = emu_taylor_tangents(cfg, n_inst, params, y0, grid, n_tan);
  const int Kobs = (K - 1) / every + 1;
  const size_t csz = 2 * (size_t)L.nc * L.nc + (size_t)L.nc * L.mc + 2 * (size_t)L.nc;
  std::vector<double> scratch(csz * (size_t)(Kobs > 1 ? Kobs - 1 : 1) * n_inst);
  int queue = 0;
  pdq::KArgs a{};
  a.cfg = *cfg;
  a.lay = L;
  a.n_inst = n_inst;
  a.K = K;
  a.mode = pdq::MODE_LML;
  a.every = every;
  a.n_tan = n_tan;
  a.average_pdfs = average;
  a.n_std = n_std;
  a.params = params;
  a.tcoeffs = tinit.data();
  a.times = grid;
  a.data = data;
  a.stds = stds;
  a.A = A.data();
  a.Q = Q.data();
  a.lml_value = value;
  a.lml_grad = grad;
  a.scratch = scratch.data();
  a.queue = &queue;
  std::vector<double> shared(L.wg_shared), ws(L.inst_size);
  for (int i = 0; i < n_inst; ++i) pdq::emu_solve_instance(a, i, shared.data(), ws.data());
  return 0;
}

// The lane-per-instance LML (pdq_lane_lml.h) on the CPU: every (instance,
// tangent group) lane in turn, with the device's slot and scratch layouts.
// tpl: tangents per lane (0 without gradient).  Returns 2 for a shape the
// device library does not instantiate.
int pdqemu_lml_lane(const pdq_config* cfg, int n_inst, const double* params, const double* y0, const double* grid,
                    int K, int every, const double* data, const double* stds, int average, double* value, double* grad,
                    int tpl) {
  const int n_tan = grad ? cfg->n_params + cfg->ode_dim : 0;
  const int N = cfg->num_derivatives + 1, D = cfg->ode_dim;
  if (cfg->factorisation != PDQ_ISOTROPIC || D != 2 || (N != 4 && N != 5)) return 2;
  if ((n_tan == 0) != (tpl == 0) || (tpl > 0 && n_tan % tpl != 0)) return 2;
  std::vector<double> A, Q;
  pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  std::vector<double> tinit = emu_taylor_tangents(cfg, n_inst, params, y0, grid, n_tan);
  const int G = tpl ? n_tan / tpl : 1;
  const int Kobs = (K - 1) / every + 1;
  const int MC = D * (1 + tpl);
  const size_t csz = (size_t)N * N + (size_t)N * MC + (size_t)N * (N + 1) / 2 + 2 * (size_t)N;
  const size_t n_lanes = (size_t)n_inst * G;
  std::vector<double> scratch(csz * (size_t)(Kobs > 1 ? Kobs - 1 : 1) * n_lanes);
  pdq::LmlArgs a{};
  a.cfg = *cfg;
  a.cfg.strategy = PDQ_FIXEDPOINT;
  a.n_inst = n_inst;
  a.K = K;
  a.every = every;
  a.n_tan = n_tan;
  a.average_pdfs = average;
  a.G = G;
  a.params = params;
  a.tcoeffs = tinit.data();
  a.times = grid;
  a.data = data;
  a.stds = stds;
  a.Q = Q.data();
  a.value = value;
  a.grad = grad;
  a.scratch = scratch.data();
  pdq::LArgs la{};
  la.cfg = a.cfg;
  la.n_inst = n_inst;
  la.K = K;
  la.Q = Q.data();
  std::vector<double> slots(64 * 2 * 256);
  for (size_t gl = 0; gl < n_lanes; ++gl) {
    const int inst = (int)(gl / G), g = (int)(gl % G);
#define PDQEMU_LML(NN, TT) pdq::lane_lml_instance<NN, 2, TT>(la, a, inst, g, slots.data(), gl, n_lanes)
    if (N == 4 && tpl == 0) PDQEMU_LML(4, 0);
    else if (N == 4 && tpl == 1) PDQEMU_LML(4, 1);
    else if (N == 4 && tpl == 2) PDQEMU_LML(4, 2);
    else if (N == 4 && tpl == 3) PDQEMU_LML(4, 3);
    else if (N == 5 && tpl == 0) PDQEMU_LML(5, 0);
    else if (N == 5 && tpl == 1) PDQEMU_LML(5, 1);
    else if (N == 5 && tpl == 2) PDQEMU_LML(5, 2);
    else if (N == 5 && tpl == 3) PDQEMU_LML(5, 3);
    else return 2;
#undef PDQEMU_LML
  }
  return 0;
}


// The reverse-mode adjoint of the lane LML (pdq_lane_adjoint.h), one lane per
// instance, `lanes` scratch columns (instances grid-stride over them as on
// the device).
int pdqemu_lml_adjoint(const pdq_config* cfg, int n_inst, const double* params, const double* y0, const double* grid,
                       int K, int every, const double* data, const double* stds, int average, double* value,
                       double* grad, int lanes) {
  const int n_tan = grad ? cfg->n_params + cfg->ode_dim : 0;
  const int N = cfg->num_derivatives + 1, D = cfg->ode_dim;
  if (cfg->factorisation != PDQ_ISOTROPIC || D != 2 || (N != 4 && N != 5) || lanes < 1) return 2;
  std::vector<double> A, Q;
  pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  std::vector<double> tinit = emu_taylor_tangents(cfg, n_inst, params, y0, grid, n_tan);
  const size_t lsz = N == 4 ? pdq::AdjLayout<4, 2>(K, every).size : pdq::AdjLayout<5, 2>(K, every).size;
  std::vector<double> scratch(lsz * (size_t)lanes, 0.0);
  pdq::LmlArgs a{};
  a.cfg = *cfg;
  a.cfg.strategy = PDQ_FIXEDPOINT;
  a.n_inst = n_inst;
  a.K = K;
  a.every = every;
  a.n_tan = n_tan;
  a.average_pdfs = average;
  a.G = 1;
  a.params = params;
  a.tcoeffs = tinit.data();
  a.times = grid;
  a.data = data;
  a.stds = stds;
  a.Q = Q.data();
  a.value = value;
  a.grad = grad;
  a.scratch = scratch.data();
  pdq::LArgs la{};
  la.cfg = a.cfg;
  la.n_inst = n_inst;
  la.K = K;
  la.Q = Q.data();
  std::vector<double> slots(64 * 2 * 256);
  for (int inst = 0; inst < n_inst; ++inst) {
    const size_t lane = (size_t)inst % lanes;
    if (N == 4) pdq::lane_adj_instance<4, 2>(la, a, inst, slots.data(), lane, (size_t)lanes);
    else pdq::lane_adj_instance<5, 2>(la, a, inst, slots.data(), lane, (size_t)lanes);
  }
  return 0;
}

// The off-grid marginal kernel (pdq_lane_offgrid.h) on the CPU, one query
// after another.  Returns 2 for a shape or strategy it does not cover.
int pdqemu_offgrid(const pdq_config* cfg, int n_inst, int K, const double* times, const double* mean,
                   const double* chol, const double* filt_mean, const double* filt_chol, const double* output_scale,
                   int n_query, const double* t_query, double* mean_out, double* chol_out) {
  const int N = cfg->num_derivatives + 1, D = cfg->ode_dim;
  if (cfg->factorisation == PDQ_BLOCKDIAG || cfg->strategy == PDQ_FIXEDPOINT) return 2;
  std::vector<double> A, Q;
  if (cfg->factorisation == PDQ_DENSE || D != 2 || (N != 4 && N != 5) || cfg->lanes_per_instance > 1) {
    // the cooperative kernel's off-grid code, one query after another
    if (cfg->factorisation == PDQ_DENSE) pdq::iwp_dense(cfg->num_derivatives, D, cfg->base_output_scale, A, Q);
    else pdq::iwp_matrices(cfg->num_derivatives, A, Q);
    pdq::OffgridCoopArgs o{};
    o.k.cfg = *cfg;
    o.k.lay = pdq::make_layout(*cfg, 0);
    o.k.n_inst = n_inst;
    o.k.K = K;
    o.k.times = times;
    o.k.A = A.data();
    o.k.Q = Q.data();
    o.n_query = n_query;
    o.t_query = t_query;
    o.mean = mean;
    o.chol = chol;
    o.filt_mean = filt_mean;
    o.filt_chol = filt_chol;
    o.output_scale = output_scale;
    o.mean_out = mean_out;
    o.chol_out = chol_out;
    std::vector<double> shared(o.k.lay.wg_shared), ws(o.k.lay.inst_size);
    for (size_t q = 0; q < (size_t)n_inst * n_query; ++q) pdq::emu_offgrid_query(o, q, shared.data(), ws.data());
    return 0;
  }
  pdq::iwp_matrices(cfg->num_derivatives, A, Q);
  pdq::OffgridArgs a{};
  a.cfg = *cfg;
  a.n_inst = n_inst;
  a.K = K;
  a.n_query = n_query;
  a.times = times;
  a.t_query = t_query;
  a.mean = mean;
  a.chol = chol;
  a.filt_mean = filt_mean;
  a.filt_chol = filt_chol;
  a.output_scale = output_scale;
  a.Q = Q.data();
  a.mean_out = mean_out;
  a.chol_out = chol_out;
  const size_t total = (size_t)n_inst * n_query;
  for (size_t q = 0; q < total; ++q) {
    if (N == 4 && cfg->strategy == PDQ_FILTER) pdq::lane_offgrid<4, 2, PDQ_FILTER>(a, q);
    else if (N == 4) pdq::lane_offgrid<4, 2, PDQ_FIXEDINTERVAL>(a, q);
    else if (cfg->strategy == PDQ_FILTER) pdq::lane_offgrid<5, 2, PDQ_FILTER>(a, q);
    else pdq::lane_offgrid<5, 2, PDQ_FIXEDINTERVAL>(a, q);
  }
  return 0;
}


// The posterior sampler (pdq_lane_sample.h) on the CPU, one (instance,
// sample) lane after another.
int pdqemu_sample(const pdq_config* cfg, int n_inst, int K, const double* term_mean, const double* term_chol,
                  const double* cA, const double* cm, const double* cL, const double* ctl, const double* cto,
                  int n_samples, const double* base, double* samples) {
  const int N = cfg->num_derivatives + 1, D = cfg->ode_dim;
  if (cfg->factorisation == PDQ_BLOCKDIAG) return 2;
  const bool lane = cfg->factorisation == PDQ_ISOTROPIC && D == 2 && (N == 4 || N == 5) && cfg->lanes_per_instance <= 1;
  pdq::SampleArgs a{};
  a.cfg = *cfg;
  a.n_inst = n_inst;
  a.K = K;
  a.n_samples = n_samples;
  a.term_mean = term_mean;
  a.term_chol = term_chol;
  a.cond_A = cA;
  a.cond_mean = cm;
  a.cond_chol = cL;
  a.cond_tl = ctl;
  a.cond_to = cto;
  a.base = base;
  a.samples = samples;
  if (!lane) {
    // the generic sampler (dense, other isotropic shapes), one sample after another
    const int dense = cfg->factorisation == PDQ_DENSE;
    const int n = dense ? N * D : N, mc = dense ? 1 : D;
    if (n * mc > pdq::kSampleGenericMax) return 2;
    std::vector<double> v(3 * pdq::kSampleGenericMax);
    for (size_t g = 0; g < (size_t)n_inst * n_samples; ++g) pdq::sample_generic(a, g, n, mc, dense, v.data(), 1);
    return 0;
  }
  for (size_t g = 0; g < (size_t)n_inst * n_samples; ++g) {
    if (N == 4) pdq::lane_sample<4, 2>(a, g);
    else pdq::lane_sample<5, 2>(a, g);
  }
  return 0;
}

}  // extern "C"
