// Block-diagonal factorisation (ssm_impl_blockdiag.py) on gfx950.
//
// Mapping: one workgroup owns one IVP instance at a time (persistent, work
// queue); each thread owns BPT of the d independent (nu+1) x (nu+1) blocks
// and keeps their means and Cholesky factors in registers.  Blocks couple
// only through the vector field (the neighbour exchange of a stencil goes
// through LDS) and through the scalar error norm / accept decision
// (workgroup reduction).  Per-dimension output scales follow
// BlockDiagNormal.residual_whitened_rms_flat (ssm_impl_blockdiag.py:285-292).
#include "pdq_args.h"
#include "pdq_device.h"

namespace pdq {



template <int N>
struct Blk {
  double m[N];
  double L[N * N];
};

// Householder R of an (M x C) column-major-in-registers matrix, LAPACK sign
// convention; A is row-major M x C, result in rows 0..C-1.
template <int M, int C>
__device__ __forceinline__ void qr_regs(double (&A)[M * C]) {
#pragma unroll
  for (int i = 0; i < C; ++i) {
    if (M - i <= 1) break;
    double ss = 0.0;
#pragma unroll
    for (int r = i + 1; r < M; ++r) ss += A[r * C + i] * A[r * C + i];
    if (ss != 0.0) {
      const double alpha = A[i * C + i];
      const double beta = -copysign(hypot(alpha, sqrt(ss)), alpha);
      const double tau = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
#pragma unroll
      for (int r = i + 1; r < M; ++r) A[r * C + i] *= scal;
#pragma unroll
      for (int j = i + 1; j < C; ++j) {
        double w = A[i * C + j];
#pragma unroll
        for (int r = i + 1; r < M; ++r) w += A[r * C + i] * A[r * C + j];
        w *= tau;
        A[i * C + j] -= w;
#pragma unroll
        for (int r = i + 1; r < M; ++r) A[r * C + j] -= A[r * C + i] * w;
      }
      A[i * C + i] = beta;
    }
  }
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#if defined(PDQ_HOST_EMU) && !defined(PDQ_NO_EMU_BLOCKSUM)
  return v;  // one thread per workgroup in the CPU build
#endif
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  const int w = threadIdx.x / 64, nw = blockDim.x / 64;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

template <int N, int BPT>
__device__ void solve_instance_bd(const BArgs& a, int inst, double* y, double* red) {
  const pdq_config& cfg = a.cfg;
  const int d = cfg.ode_dim, K = a.K;
  const double* ts = a.times + (size_t)inst * K;
  const double* prm = a.params + (size_t)inst * cfg.n_params;
  double A1[N * N], Q1[N * N];
#pragma unroll
  for (int i = 0; i < N * N; ++i) {
    A1[i] = a.A1[i];
    Q1[i] = a.Q1[i];
  }
  int blk[BPT];
  bool act[BPT];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    blk[q] = threadIdx.x + q * blockDim.x;
    act[q] = blk[q] < d;
  }
  Blk<N> sf[BPT], fr[BPT], pr[BPT];
  double os_sf[BPT], os_pr[BPT], b_pr[BPT], b_sf[BPT];
  double t_sf, t_fr, t_pr, ns_sf, ns_fr, ns_pr;
  // prior init (BlockDiagNormal.from_mean_and_std)
  const double* tc = a.tcoeffs + (size_t)inst * N * d;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
#pragma unroll
    for (int i = 0; i < N; ++i) sf[q].m[i] = act[q] ? tc[i * d + blk[q]] : 0.0;
#pragma unroll
    for (int i = 0; i < N * N; ++i) sf[q].L[i] = (i % (N + 1) == 0) ? cfg.init_std : 0.0;
    os_sf[q] = 1.0;
    b_sf[q] = 0.0;
  }
  t_sf = ts[0];
  ns_sf = 0.0;
  const pdq_solution& o = a.out;
  const int ms = d * N, cs = d * N * N;
  auto write = [&](int k, double t, const Blk<N>* x, const double* os, double nsteps) {
    const size_t row = (size_t)inst * K + k;
    if (threadIdx.x == 0) {
      if (o.t) o.t[row] = t;
      if (o.num_steps && k > 0) o.num_steps[(size_t)inst * (K - 1) + k - 1] = (int64_t)nsteps;
    }
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      if (!act[q]) continue;
      if (o.mean)
#pragma unroll
        for (int i = 0; i < N; ++i) o.mean[row * ms + blk[q] * N + i] = x[q].m[i];
      if (o.chol)
#pragma unroll
        for (int i = 0; i < N * N; ++i) o.chol[row * cs + (size_t)blk[q] * N * N + i] = x[q].L[i];
      if (o.output_scale && cfg.solver == PDQ_SOLVER_DYNAMIC) o.output_scale[row * d + blk[q]] = os[q];
    }
  };
  double one_os[BPT];
#pragma unroll
  for (int q = 0; q < BPT; ++q) one_os[q] = 1.0;
  write(0, t_sf, sf, one_os, 0.0);
#pragma unroll
  for (int q = 0; q < BPT; ++q) fr[q] = sf[q];
  t_fr = t_sf;
  ns_fr = ns_sf;

  double P[N], PI[N];
  auto precon = [&](double dt) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int pw = N - 1 - i;
      const double s = factorial_d(pw);
      P[i] = pow(dt, (double)pw) / s;
      PI[i] = pow(dt, -(double)pw) * s;
    }
  };
  // transition.marginalise(x) with output scale os (ssm_impl_blockdiag.py:28-43)
  auto marg_prior = [&](double dt, double osq, const Blk<N>& x, Blk<N>& y) {
    const double s = sqrt(fabs(dt)) * (cfg.base_output_scale * osq);
    double S[2 * N * N];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int col = 0; col < N; ++col) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) acc += A1[col * N + k] * (fabs(PI[k]) * x.L[k * N + r]);
        S[r * N + col] = acc;
        S[(N + r) * N + col] = s * Q1[col * N + r];
      }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += A1[i * N + k] * (PI[k] * x.m[k]);
      y.m[i] = P[i] * (acc + 0.0);
    }
    qr_regs<2 * N, N>(S);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) y.L[i * N + j] = j <= i ? fabs(P[i]) * S[j * N + i] : 0.0;
  };
  // fx.marginalise(u) for the ts0 selector (derivative 1): observed scalar
  auto marg_fx = [&](double b, const Blk<N>& u, double& om, double& oL) {
    double col[N + 1];
#pragma unroll
    for (int r = 0; r < N; ++r) col[r] = u.L[1 * N + r];
    col[N] = cfg.damp;
    qr_regs<N + 1, 1>(col);
    oL = col[0];
    om = u.m[1] + b;
  };
  // fx.revert(pred) + apply_flat(0)
  auto update = [&](double b, const Blk<N>& pm, Blk<N>& x) {
    constexpr int M = N + 1;
    double R[M * M];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
      for (int col = 0; col < M; ++col) {
        double v;
        if (r == 0) v = col == 0 ? cfg.damp : 0.0;
        else v = col == 0 ? pm.L[1 * N + (r - 1)] : pm.L[(col - 1) * N + (r - 1)];
        R[r * M + col] = v;
      }
    qr_regs<M, M>(R);
    const double mo = pm.m[1] + b;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double g = R[1 + i] / R[0];
      x.m[i] = pm.m[i] - g * mo;
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) x.L[i * N + j] = j <= i ? R[(1 + j) * M + 1 + i] : 0.0;
  };
  // linearize: b = -f(u(t)) with the neighbour exchange through LDS
  auto linearize = [&](const Blk<N>* u, double t, double* b) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; ++q)
      if (act[q]) y[blk[q]] = u[q].m[0];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; ++q) b[q] = act[q] ? -vf_component(cfg.vector_field, d, y, prm, t, blk[q]) : 0.0;
  };
  // one step attempt: proposed <- step(step_from, dt); returns error power
  auto step = [&](double dt, bool estimate) -> double {
    precon(dt);
    const double tn = t_sf + dt;
    double ss = 0.0;
    if (cfg.solver == PDQ_SOLVER_DYNAMIC) {
      Blk<N> u[BPT];
      const double s1 = sqrt(fabs(dt)) * (cfg.base_output_scale * 1.0);
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          double acc = 0.0;
#pragma unroll
          for (int k = 0; k < N; ++k) acc += A1[i * N + k] * (PI[k] * sf[q].m[k]);
          u[q].m[i] = P[i] * (acc + 0.0);
        }
#pragma unroll
        for (int i = 0; i < N * N; ++i) u[q].L[i] = fabs(P[i / N]) * (s1 * Q1[i]);
      }
      linearize(u, tn, b_pr);
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        double om, oL;
        marg_fx(b_pr[q], u[q], om, oL);
        const double sigma = fabs((0.0 - om) / oL) / sqrt(1.0);
        os_pr[q] = sigma;
        Blk<N> pm;
        marg_prior(dt, sigma, sf[q], pm);
        update(b_pr[q], pm, pr[q]);
        if (estimate && act[q]) {
          const double e = sqrt((sigma * oL) * (sigma * oL)) * pow(dt, 1.0) / 1.0;
          const double ref = fmax(fabs(sf[q].m[0]), fabs(pr[q].m[0]));
          const double r = e / (cfg.atol + cfg.rtol * fabs(ref));
          ss += r * r;
        }
      }
    } else {
      Blk<N> pm[BPT];
#pragma unroll
      for (int q = 0; q < BPT; ++q) marg_prior(dt, 1.0, sf[q], pm[q]);
      linearize(pm, tn, b_pr);
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        update(b_pr[q], pm[q], pr[q]);
        os_pr[q] = 1.0;
      }
      if (estimate) {
        const double s1 = sqrt(fabs(dt)) * (cfg.base_output_scale * 1.0);
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
          Blk<N> u;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += A1[i * N + k] * (PI[k] * sf[q].m[k]);
            u.m[i] = P[i] * (acc + 0.0);
          }
#pragma unroll
          for (int i = 0; i < N * N; ++i) u.L[i] = fabs(P[i / N]) * (s1 * Q1[i]);
          double om, oL;
          marg_fx(b_pr[q], u, om, oL);
          const double se = fabs((0.0 - om) / oL) / sqrt(1.0);
          if (act[q]) {
            const double e = sqrt((se * oL) * (se * oL)) * pow(dt, 1.0) / 1.0;
            const double ref = fmax(fabs(sf[q].m[0]), fabs(pr[q].m[0]));
            const double r = e / (cfg.atol + cfg.rtol * fabs(ref));
            ss += r * r;
          }
        }
      }
    }
    t_pr = tn;
    ns_pr = ns_sf + 1.0;
    if (!estimate) return 2.0;
    const double tot = block_sum(ss, red);
    const double norm = sqrt(tot) / sqrt((double)d);
    return pow(norm, -1.0 / (double)N);
  };

  long attempts = 0, accepted = 0;
  int status = PDQ_STATUS_OK;
  if (a.mode == BMODE_FIXED_GRID) {
    for (int k = 1; k < K; ++k) {
      step(ts[k] - ts[k - 1], false);
      ++attempts;
      ++accepted;
      write(k, t_pr, pr, os_pr, ns_pr);
#pragma unroll
      for (int q = 0; q < BPT; ++q) {
        sf[q] = pr[q];
        os_sf[q] = os_pr[q];
        b_sf[q] = b_pr[q];
      }
      t_sf = t_pr;
      ns_sf = ns_pr;
    }
  } else {
    double dt = cfg.dt0_per_instance ? cfg.dt0_per_instance[inst] : cfg.dt0, cstate = 1.0;
    for (int k = 1; k < K && status == PDQ_STATUS_OK; ++k) {
      const double t1 = ts[k];
      bool cont = true;
      while (cont) {
        if (t_sf + cfg.eps < t1) {
          double ep = 0.9;
          int att = 0;
          while (ep < 1.0) {
            double dtc = dt;
            if (cfg.clip_dt) dtc = fmin(dtc, t1 - t_sf);
            ep = step(dtc, true);
            // controllers.py:66-84 / :24-63
            if (cfg.control == PDQ_CONTROL_INTEGRAL) {
              dt = fmax(cfg.factor_min, fmin(cfg.safety * ep, cfg.factor_max)) * dtc;
            } else {
              const double gi = pow(ep, cfg.exponent_integral);
              const double gp = pow(ep / cstate, cfg.exponent_proportional);
              const double sfac = fmax(cfg.factor_min, fmin(cfg.safety * gi * gp, cfg.factor_max));
              if (ep >= 1.0) cstate = ep;
              dt = sfac * dtc;
            }
            ++att;
            ++attempts;
            if (!isfinite(ep) || !isfinite(dt)) {
              status = PDQ_STATUS_NONFINITE;
              break;
            }
            if (ep < 1.0 && (dt == 0.0 || t_sf + dt == t_sf)) {
              status = PDQ_STATUS_DT_UNDERFLOW;
              break;
            }
            if (ep >= 1.0 && t_pr == t_sf) {
              status = PDQ_STATUS_DT_UNDERFLOW;  // accepted without advancing t
              break;
            }
            if ((cfg.max_attempts > 0 && att >= cfg.max_attempts && ep < 1.0) || attempts >= kMaxTotalAttempts) {
              status = PDQ_STATUS_MAX_ATTEMPTS;
              break;
            }
          }
          if (status != PDQ_STATUS_OK) break;
          ++accepted;
#pragma unroll
          for (int q = 0; q < BPT; ++q) {
            fr[q] = sf[q];
            sf[q] = pr[q];
            os_sf[q] = os_pr[q];
            b_sf[q] = b_pr[q];
          }
          t_fr = t_sf;
          ns_fr = ns_sf;
          t_sf = t_pr;
          ns_sf = ns_pr;
        }
        if (t_sf + cfg.eps < t1) {
          cont = true;
        } else if (t_sf > t1 + cfg.eps) {
          // strategy_filter.interpolate_fwd: predict from interp_from with
          // interp_to's output scale (solvers.py:205-269)
          precon(t1 - t_fr);
#pragma unroll
          for (int q = 0; q < BPT; ++q) {
            Blk<N> y2;
            marg_prior(t1 - t_fr, os_sf[q], fr[q], y2);
            fr[q] = y2;
          }
          t_fr = t1;
          write(k, t1, fr, os_sf, ns_sf);
          cont = t_sf + cfg.eps < t1;
        } else {
          write(k, t_sf, sf, os_sf, ns_sf);
#pragma unroll
          for (int q = 0; q < BPT; ++q) fr[q] = sf[q];
          t_fr = t_sf;
          ns_fr = ns_sf;
          cont = false;
        }
      }
    }
  }
  (void)ns_fr;
  (void)b_sf;
  if (threadIdx.x == 0) {
    if (o.attempts) o.attempts[inst] = attempts;
    if (o.accepted) o.accepted[inst] = accepted;
    if (o.status) o.status[inst] = status;
    if (o.output_scale && cfg.solver == PDQ_SOLVER)
      for (int k = 0; k < K - 1; ++k)
        for (int j = 0; j < d; ++j) o.output_scale[((size_t)inst * K + k) * d + j] = 1.0;
  }
}

#ifndef PDQ_HOST_EMU
template <int N, int BPT>
__global__ void __launch_bounds__(1024) blockdiag_kernel(BArgs a) {
  extern __shared__ double lds[];
  double* y = lds;                    // d entries: stencil exchange
  double* red = lds + a.cfg.ode_dim;  // 16 entries: reductions
  __shared__ int s_inst;
  for (;;) {
    if (threadIdx.x == 0) s_inst = atomicAdd(a.queue, 1);
    __syncthreads();
    const int inst = s_inst;
    __syncthreads();
    if (inst >= a.n_inst) break;
    solve_instance_bd<N, BPT>(a, inst, y, red);
  }
}


__global__ void zero_queue_bd(int* p) {
  if (threadIdx.x == 0) p[0] = 0;
}

template <int N, int BPT>
int launch_bd_nb(BArgs& a, int threads, hipStream_t st) {
  const size_t lds = sizeof(double) * (size_t)(a.cfg.ode_dim + 16);
  auto kern = blockdiag_kernel<N, BPT>;
  int per_cu = 0;
  PDQ_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds));
  if (per_cu < 1) per_cu = 1;
  const long cap = (long)per_cu * num_cus();
  const int grid = (int)std::max<long>(1, std::min<long>(a.n_inst, cap));
  hipLaunchKernelGGL(zero_queue_bd, dim3(1), dim3(64), 0, st, a.queue);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, st, a);
  PDQ_HIP_TRY(hipGetLastError());
  return PDQ_OK;
}

template <int N>
int launch_bd_n(BArgs& a, hipStream_t st) {
  const int d = a.cfg.ode_dim;
  int threads = ((d + 63) / 64) * 64;
  int bpt = 1;
  while (threads > 1024) {
    bpt *= 2;
    threads = (((d + bpt - 1) / bpt + 63) / 64) * 64;
  }
  switch (bpt) {
    case 1: return launch_bd_nb<N, 1>(a, threads, st);
    case 2: return launch_bd_nb<N, 2>(a, threads, st);
    case 4: return launch_bd_nb<N, 4>(a, threads, st);
    case 8: return launch_bd_nb<N, 8>(a, threads, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "blockdiag supports ode_dim <= 8192");
}

int launch_bd(BArgs& a, hipStream_t st) {
  switch (a.cfg.num_derivatives + 1) {
    case 2: return launch_bd_n<2>(a, st);
    case 3: return launch_bd_n<3>(a, st);
    case 4: return launch_bd_n<4>(a, st);
    case 5: return launch_bd_n<5>(a, st);
    case 6: return launch_bd_n<6>(a, st);
  }
  return host_error(PDQ_ERR_UNSUPPORTED, "blockdiag supports num_derivatives in [1, 5]");
}


#else
// CPU build: one thread owns all d blocks (BPT >= d); y holds d doubles.
template <int N, int BPT>
void emu_solve_instance_bd(const BArgs& a, int inst, double* y, double* red) {
  threadIdx.x = 0;
  blockDim.x = 1;
  solve_instance_bd<N, BPT>(a, inst, y, red);
}
#endif
}  // namespace pdq
