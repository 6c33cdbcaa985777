/*
 * pdq.h -- C ABI of the MI355X (gfx950) probabilistic IVP solver path.
 *
 * Every entry point takes device pointers (hipMalloc'ed or torch-owned),
 * plain sizes and a hipStream_t passed as void*.  Calls are stream-ordered;
 * no pointer is retained after the call's kernels complete.  Return value 0
 * means success; a non-zero PDQ_ERR_* code leaves a message readable with
 * pdq_last_error().  There is no CPU fallback: without a HIP device every
 * compute entry point returns PDQ_ERR_NO_DEVICE.
 *
 * Flat layouts (row-major) follow the reference's flattened normals:
 *   dense:      mean n = (nu+1)*d (derivative-major), chol n x n
 *   isotropic:  mean (nu+1) x d,                        chol (nu+1) x (nu+1)
 *   blockdiag:  mean d x (nu+1),                        chol d x (nu+1) x (nu+1)
 * (probdiffeq/_probdiffeq/ssm_impl_dense.py:108-129,
 *  ssm_impl_isotropic.py:146-178, ssm_impl_blockdiag.py:215-253)
 */
#ifndef PDQ_H_
#define PDQ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* factorisation: state_space_model_{dense,isotropic,blockdiag}
 * (ssm_impl_dense.py:455, ssm_impl_isotropic.py:406, ssm_impl_blockdiag.py:460) */
enum { PDQ_DENSE = 0, PDQ_ISOTROPIC = 1, PDQ_BLOCKDIAG = 2 };
/* constraint_ode_ts0 / constraint_ode_ts1 (ssm_impl_api.py:520-558) */
enum { PDQ_TS0 = 0, PDQ_TS1 = 1 };
/* strategy_filter / _smoother_fixedpoint / _smoother_fixedinterval
 * (estimators_and_losses.py:347, 473, 594) */
enum { PDQ_FILTER = 0, PDQ_FIXEDPOINT = 1, PDQ_FIXEDINTERVAL = 2 };
/* solver / solver_dynamic / solver_mle (solvers.py:636, 483, 318) */
enum { PDQ_SOLVER = 0, PDQ_SOLVER_DYNAMIC = 1, PDQ_SOLVER_MLE = 2 };
/* error_norm_scale_then_rms / _rms_then_scale (solvers.py:770-811) */
enum { PDQ_NORM_SCALE_THEN_RMS = 0, PDQ_NORM_RMS_THEN_SCALE = 1 };
/* control_integral / control_proportional_integral (controllers.py:24-84) */
enum { PDQ_CONTROL_INTEGRAL = 0, PDQ_CONTROL_PI = 1 };
/* error_residual_std (solvers.py:850-996), error_state_std (:999-1098) */
enum { PDQ_ERROR_RESIDUAL_STD = 0, PDQ_ERROR_STATE_STD = 1 };
/* Priors: integrated Wiener process (ssm_impl_*.py prior_wiener_integrated)
 * and the dense exponential-integrator prior (ssm_impl_dense.py:392-629). */
enum { PDQ_PRIOR_IWP = 0, PDQ_PRIOR_EXPONENTIAL = 1 };
/* LML gradient modes (pdq_config.lml_gradient) */
enum { PDQ_GRAD_FORWARD = 0, PDQ_GRAD_REVERSE = 1 };
/* LML gain solves (pdq_config.lml_solve) */
enum { PDQ_LML_SOLVE_LSTSQ = 0, PDQ_LML_SOLVE_TRIU = 1 };
/* Jacobian handlers of the ts1 linearisation (pdq_config.jacobian): the exact
 * trace / diagonal (jacobian_materialize, jacobians.py:78-113), the
 * reverse-mode Monte Carlo estimate (jacobian_monte_carlo_rev, :195-271, the
 * reference's default for ode()) or the forward-mode one
 * (jacobian_monte_carlo_fwd, :116-192).  The two estimates of the trace are
 * the same v^T J v; of the diagonal, the reverse mode takes (v^T J)_k v_k and
 * the forward mode (J v)_k v_k.  The dense factorisation always takes the
 * exact Jacobian (materialize_dense). */
enum { PDQ_JAC_EXACT = 0, PDQ_JAC_MONTE_CARLO_REV = 1, PDQ_JAC_MONTE_CARLO_FWD = 2 };
/* device vector-field registry (problems.py:283-295 convention u' = f(u, t)) */
enum {
  PDQ_VF_LOTKA_VOLTERRA = 0, /* params (a, b, c, d), d = 2 */
  PDQ_VF_VAN_DER_POL = 1,    /* params (mu), d = 2 */
  PDQ_VF_BURGERS = 2,        /* params (viscosity, dx), periodic grid of d points */
  PDQ_VF_LINEAR = 3,         /* params lambda (d): f = lambda * u + t */
  PDQ_VF_LOGISTIC = 4        /* params (r): f = r u (1 - u) */
};

/* Per-instance status (the reference has no runtime error channel). */
enum { PDQ_STATUS_OK = 0, PDQ_STATUS_MAX_ATTEMPTS = 1, PDQ_STATUS_NONFINITE = 2, PDQ_STATUS_DT_UNDERFLOW = 3 };

/* Return codes. */
enum {
  PDQ_OK = 0,
  PDQ_ERR_INVALID = 1,
  PDQ_ERR_UNSUPPORTED = 2,
  PDQ_ERR_NO_DEVICE = 3,
  PDQ_ERR_HIP = 4,
  PDQ_ERR_INSTANCE = 5 /* at least one instance ended with a non-OK status */
};

/* POD mirror of the reference's configuration objects.  Defaults of the
 * reference: safety 0.95, factor_min 0.2, factor_max 10 (controllers.py:
 * 478), dt0 0.1, eps 1e-8, damp 0 (solvers_via_adaptive_steps.py:101). */
typedef struct pdq_config {
  int32_t factorisation;
  int32_t num_derivatives; /* nu; the prior has nu+1 Taylor coefficients */
  int32_t ode_dim;         /* d */
  int32_t constraint;
  int32_t strategy;
  int32_t solver;
  int32_t error_norm;
  int32_t control;
  int32_t vector_field;
  int32_t n_params;        /* vector-field parameters per instance */
  int32_t clip_dt;
  int32_t max_attempts;    /* per accepted step; 0 = unbounded (reference) */
  double rtol, atol, dt0, eps, damp;
  double safety, factor_min, factor_max, exponent_integral, exponent_proportional;
  double base_output_scale;   /* prior_wiener_integrated(output_scale=...) */
  double init_std;            /* 0: is_exact=True; else inexact_eps */
  int32_t correct_asymptotic_underconfidence; /* solver_mle */
  int32_t lanes_per_instance; /* 0: automatic; 1: one lane per instance (isotropic, nu = 3 or 4, d = 2);
                                 4, 8, 16, 32, 64: cooperative kernel with that many lanes per instance */
  int32_t re_linearize_after_calibration; /* solver_dynamic(re_linearize_after_calibration=True),
                                             solvers.py:576-582 */
  /* Error estimator (solvers.py:850-1098): PDQ_ERROR_RESIDUAL_STD or
   * PDQ_ERROR_STATE_STD, and their options re_linearize_before_error,
   * derivative_idx (state std only) and error_per_unit_step. */
  int32_t error_estimator;
  int32_t error_re_linearize;
  int32_t error_derivative_idx;
  int32_t error_per_unit_step;
  /* Prior.  PDQ_PRIOR_EXPONENTIAL (dense factorisation, (nu+1) d <= 16,
   * solves only): prior_exponential / prior_ornstein_uhlenbeck_integrated /
   * prior_matern (ssm_impl_dense.py:516-629, ssm_impl_api.py:381-510), the
   * linear SDE dx = A x dt + B dW with A the IWP's shift matrix whose last d
   * rows are prior_drift and B = kron(e_last, base_output_scale I); each
   * transition is exp_gram_cholesky(pade_and_legendre_9) of the
   * preconditioned A dt and B (util/gram_util.py:17-77).  prior_drift is a
   * device pointer to d x (nu+1) d doubles (row-major, derivative-major
   * columns), shared by the batch. */
  int32_t prior;
  const double* prior_drift;
  /* Optional initial step per instance (device, n_inst doubles), as a vmap of
   * the reference's solve over per-instance dt0 (ivpsolve.dt0 /
   * dt0_adaptive, stepsize_initialisers.py:7-67); NULL: dt0 for every
   * instance. */
  const double* dt0_per_instance;
  /* loss_lml_timeseries(tcoeff_index=) (estimators_and_losses.py:53-105):
   * the Taylor coefficient the LML's data observe, in [0, num_derivatives];
   * read by pdq_lml_value_and_grad only (0: the solution itself). */
  int32_t lml_tcoeff_index;
  /* Optional initial standard deviation per covariance row (device, n
   * doubles: dense (nu+1) d derivative-major, isotropic nu+1), shared by the
   * instances: prior_wiener_integrated's element-wise is_exact and
   * diffuse_derivatives (ssm_impl_isotropic.py:409-447, ssm_impl_dense.py:
   * 458-500).  NULL: init_std on every row. */
  const double* init_std_rows;
  /* pdq_lml_value_and_grad's gradient: PDQ_GRAD_FORWARD (forward-mode tangents, one lane per
   * direction) or PDQ_GRAD_REVERSE (the reverse-mode adjoint of the forward pass and the reverse
   * scan, one lane per instance, with checkpoints every 50 grid steps and the forward factors
   * recomputed: isotropic nu = 3, 4 with d = 2, ts0, lml_tcoeff_index 0). */
  int32_t lml_gradient;
  /* loss_lml_timeseries(solve_triu=) (estimators_and_losses.py:53-54): the solve behind the LML's
   * gains, PDQ_LML_SOLVE_LSTSQ (linalg.lstsq_svd, the reference's default: the minimum-norm
   * solution, singular values at or below eps k s_max dropped) or PDQ_LML_SOLVE_TRIU
   * (linalg.solve_triu).  With one observed row per node (isotropic) the two differ only for an
   * exactly zero factor, where the log-density is not finite either way. */
  int32_t lml_solve;
  /* ts1 Jacobian handler (PDQ_JAC_*), and for the Monte Carlo handlers the number of Rademacher
   * probes (1..64) and the seed of their stream.  Each probe entry is a sign drawn by a
   * counter-based generator from (seed, set, probe, coordinate), with `set` the number of
   * accepted steps before the linearisation (twice that for solver_dynamic with
   * re_linearize_after_calibration, whose second linearisation takes set + 1): a rejected
   * attempt's retry reuses its probes, as the reference's PRNG key, carried in the solution, is
   * discarded with the rejected proposal.  error_re_linearize is refused with these handlers
   * (the reference draws that linearisation's probes from the error estimator's own key).  The stream is this library's counter-based
   * generator, not JAX's threefry: values follow the same estimator, not the same draws. */
  int32_t jacobian;
  int32_t jac_probes;
  uint32_t jac_seed;
} pdq_config;

/* Caller-allocated device outputs; NULL members are not written.
 * K = number of save_at points (or grid points). */
typedef struct pdq_solution {
  double* t;            /* n_inst x K */
  double* mean;         /* n_inst x K x mean_size      (u.mean) */
  double* chol;         /* n_inst x K x chol_size      (u.cholesky) */
  double* output_scale; /* n_inst x K x nblk (solver_dynamic; K-1 rows otherwise) */
  int64_t* num_steps;   /* n_inst x (K-1) accepted steps up to each point */
  /* smoothers: solution_full (SmoothingSolution, estimators_and_losses.py:277) */
  double* filt_mean;    /* n_inst x K x mean_size */
  double* filt_chol;    /* n_inst x K x chol_size */
  double* cond_A;       /* n_inst x (K-1) x nblk*n*n  backward conditionals */
  double* cond_mean;    /* n_inst x (K-1) x mean_size */
  double* cond_chol;    /* n_inst x (K-1) x chol_size */
  double* cond_tl;      /* n_inst x (K-1) x nblk*n    to_latent */
  double* cond_to;      /* n_inst x (K-1) x nblk*n    to_observed */
  double* term_mean;    /* n_inst x mean_size         full_posterior.marginal */
  double* term_chol;    /* n_inst x chol_size */
  /* statistics (the reference counts accepted steps only) */
  int64_t* attempts;    /* n_inst */
  int64_t* accepted;    /* n_inst */
  int32_t* status;      /* n_inst, PDQ_STATUS_* */
} pdq_solution;

int32_t pdq_version(void);

/* Initialise the library on the current device (idempotent). */
int32_t pdq_init(void);

/* Copies the last error message of the calling thread into buf. */
size_t pdq_last_error(char* buf, size_t size);

/* utilities.system_matrices_1d_iwp (utilities.py:57-71): the latent IWP
 * transition A (flipped Pascal) and the lower Cholesky factor Q of the
 * flipped Hilbert matrix, (nu+1) x (nu+1) row-major each, into host memory.
 * Host only: needs no device. */
int32_t pdq_system_matrices_1d_iwp(int32_t num_derivatives, double* A, double* Q);

/* Replaces ivpsolve.solve_adaptive_save_at(solver=, error=, control=,
 * clip_dt=)(prior, save_at, atol, rtol, dt0, eps, damp)
 * (probdiffeq/_ivpsolve/solvers_via_adaptive_steps.py:46-148), batched over
 * n_inst instances (the reference batches with jax.vmap).
 *   params:  n_inst x cfg->n_params
 *   tcoeffs: n_inst x (nu+1) x d  Taylor coefficients (prior mean)
 *   save_at: n_inst x K           increasing checkpoints, save_at[0] = t0
 * Each instance owns its accept/reject loop inside one persistent launch. */
int32_t pdq_solve_adaptive_save_at(const pdq_config* cfg, int32_t n_inst, const double* params,
                                   const double* tcoeffs, const double* save_at, int32_t K,
                                   const pdq_solution* out, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): pdq_solve_adaptive_save_at with the same arguments and results,
 * but it calls no hipMalloc, hipFree or other allocating routine.  What the
 * allocating entry point allocates (the IWP matrices, the work-queue counter)
 * lives in a fixed 16 640-byte device array per GPU, held until the process
 * ends; its address is looked up by pdq_init or by the first call on a device.
 * Limit: state dimension n <= 32 (dense (nu+1) d, isotropic nu+1), the
 * kernels' own limit; larger n returns PDQ_ERR_UNSUPPORTED.  Every call on a
 * GPU rewrites that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
int32_t pdq_solve_adaptive_save_at_noalloc(const pdq_config* cfg, int32_t n_inst, const double* params,
                                           const double* tcoeffs, const double* save_at, int32_t K,
                                           const pdq_solution* out, void* stream);

/* Replaces ivpsolve.solve_fixed_grid(solver=)(prior, grid=, damp=)
 * (probdiffeq/_ivpsolve/solvers_via_fixed_steps.py:11-34).
 *   grid: n_inst x K */
int32_t pdq_solve_fixed_grid(const pdq_config* cfg, int32_t n_inst, const double* params,
                             const double* tcoeffs, const double* grid, int32_t K,
                             const pdq_solution* out, void* stream);

/* Replaces the parameter-to-loss map of examples/E2_parameter_estimation_optax.py
 * (loss_lml_timeseries, estimators_and_losses.py:53-105, evaluated on the
 * fixed-point smoother's Markov sequence, :473-591) together with its
 * gradient (jax.value_and_grad in the example).
 * The prior mean is the Taylor expansion of y0 (jetexpand, num = nu); the
 * forward pass steps the fixed grid and checkpoints every `every` points.
 *   params: n_inst x n_params, y0: n_inst x d, grid: K_grid (shared),
 *   data:   n_inst x K_obs x d  (K_obs = (K_grid-1)/every + 1)
 *   stds:   n_inst x K_obs x n_std (n_std = 1 or the per-block count)
 *   value:  n_inst, grad: n_inst x (n_params + d) (may be NULL) */
int32_t pdq_lml_value_and_grad(const pdq_config* cfg, int32_t n_inst, const double* params,
                               const double* y0, const double* grid, int32_t K_grid, int32_t every,
                               const double* data, const double* stds, int32_t n_std,
                               int32_t average_pdfs, double* value, double* grad, void* stream);

/* Replaces jetexpand_ode_padded_scan(num=nu)(vf, (y0,), t=t0)
 * (probdiffeq/_probdiffeq/jet_expansion_algorithms.py:49-103) for the
 * registered vector fields.  out: n_inst x (num+1) x d derivatives. */
int32_t pdq_taylor_coefficients(int32_t vector_field, int32_t d, int32_t n_params, int32_t n_inst,
                                const double* params, const double* y0, double t0, int32_t num,
                                double* out, void* stream);

/* Replaces ProbabilisticSolver.offgrid_marginals(t, solution=)
 * (probdiffeq/_probdiffeq/solvers.py:149-203; the strategies' parts are
 * estimators_and_losses.py:403-414 and 674-707) for a batch of solutions
 * and n_query times per instance.  The solution is what a solve with the
 * same configuration wrote: t (as `times`), mean, chol, output_scale
 * (K rows) and, for the fixed-interval smoother, filt_mean / filt_chol.
 *   t_query:  n_inst x n_query, each strictly inside (times[0], times[K-1])
 *             and off the grid points (as the reference requires)
 *   mean_out: n_inst x n_query x mean_size, chol_out: ... x chol_size
 * Filter and fixed-interval smoother; the fixed-point smoother is refused
 * (the reference raises NotImplementedError).  Isotropic nu = 3, 4, d = 2 on
 * the lane kernel; dense, the other isotropic shapes and exponential priors
 * on the cooperative kernel's primitives. */
int32_t pdq_offgrid_marginals(const pdq_config* cfg, int32_t n_inst, int32_t K, const double* times,
                              const pdq_solution* solution, int32_t n_query, const double* t_query,
                              double* mean_out, double* chol_out, void* stream);

/* Replaces loss_lml_terminal_values(tcoeff_index=idx)(u, marginals=, std=)
 * (probdiffeq/_probdiffeq/estimators_and_losses.py:20-50) for a batch of
 * isotropic or dense marginals (e.g. the terminal values of a solve):
 *   mean: n_inst x mean_size, chol: n_inst x chol_size,
 *   u: n_inst x d data, std: the observation noise, shaped as
 *   marginals.std[idx] is: one scalar per instance (isotropic) or d per
 *   instance (dense, (nu+1) d + d) d <= 320),
 *   value: n_inst log-densities. */
int32_t pdq_lml_terminal_values(const pdq_config* cfg, int32_t n_inst, const double* mean, const double* chol,
                                const double* u, const double* std, int32_t tcoeff_index, double* value,
                                void* stream);

/* Replaces MarkovSequence.sample(key, shape=(n_samples,))
 * (probdiffeq/_probdiffeq/estimators_and_losses.py:233-273) on a smoother's
 * posterior (solution.posterior: term_mean / term_chol and the K-1 backward
 * conditionals cond_*, as a solve with this configuration wrote them).  The
 * standard normal draws are an input instead of a JAX key:
 *   base:    n_inst x n_samples x K x n, n = nu+1 (isotropic: sample_flat
 *            draws one (nu+1)-vector per row for all d columns) or (nu+1) d
 *            (dense); row K-1 samples the terminal marginal, row k the
 *            conditional k
 *   samples: n_inst x n_samples x K x mean_size
 * Smoothers only (a filter's solution is no Markov sequence).  Dense and
 * isotropic with (nu+1) d <= 32 (isotropic nu = 3, 4, d = 2 on the lane
 * sampler, the other shapes on a generic one). */
int32_t pdq_sample(const pdq_config* cfg, int32_t n_inst, int32_t K, const pdq_solution* solution,
                   int32_t n_samples, const double* base, double* samples, void* stream);

/* Sizes of the flat layouts above for a configuration. */
int32_t pdq_mean_size(const pdq_config* cfg);
int32_t pdq_chol_size(const pdq_config* cfg);

/* Lanes per instance the autotuner picked for this configuration and batch
 * size (0 if not yet tuned). */
int32_t pdq_tuned_lanes(const pdq_config* cfg, int32_t n_inst);

/* Records the lanes-per-instance a timing sweep picked for this
 * configuration, batch-size bucket (next power of two) and mode (0 save_at,
 * 1 fixed grid, 2 LML); later calls with lanes_per_instance = 0 use it.
 * lanes: 1 (lane-per-instance kernel) or 4, 8, 16, 32, 64 (cooperative).
 * No counterpart in the reference (XLA picks its own launch shapes); the
 * sweep itself is probdiffeq_amd/autotune.py. */
int32_t pdq_set_tuned_lanes(const pdq_config* cfg, int32_t n_inst, int32_t mode, int32_t lanes);

#ifdef __cplusplus
}
#endif

#endif /* PDQ_H_ */
