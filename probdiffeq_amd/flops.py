"""Algorithmic FLOP model of the solver path (the reference's formulation).

Counts follow the operations the reference performs per call (LAPACK
R-only Householder QR(m, k) = 2 m k^2 - 2/3 k^3 for m >= k, GEMM 2abc,
triangular solve k^2 r, selector products H L counted as dense GEMMs), so
that ``achieved TFLOP/s = model FLOPs / measured kernel time`` compares
the kernels with the chip's FP64 peak on the work the algorithm requires,
not on what an implementation chooses to execute.

nc = covariance dimension (dense (nu+1) d, isotropic and per block nu+1),
mc = mean columns (dense 1, isotropic d, blockdiag 1), ko = observation
rows (dense d, isotropic 1, blockdiag 1 per block), nb = blocks (blockdiag
d, otherwise 1).
"""

from __future__ import annotations


def qr(m: int, k: int) -> float:
    """R-only Householder QR of an m x k matrix (LAPACK dgeqr2 count)."""
    if m < k:
        m, k = k, m
    return 2.0 * m * k * k - 2.0 / 3.0 * k**3


def gemm(a: int, b: int, c: int) -> float:
    return 2.0 * a * b * c


def trsm(k: int, r: int) -> float:
    return float(k * k * r)


def dims(cfg, n_tan: int = 0):
    """(nc, mc, ko, nb); with n_tan forward-mode directions every mean column carries n_tan
    tangent columns beside it (mc grows by the factor 1 + n_tan, the factors do not)."""
    N, d = cfg.num_derivatives + 1, cfg.ode_dim
    if cfg.factorisation == "dense":
        return N * d, 1 + n_tan, d, 1
    if cfg.factorisation == "isotropic":
        return N, d * (1 + n_tan), 1, 1
    return N, 1 + n_tan, 1, d


def _marginalise(nc, mc, rows_out):
    # (A chol)^T, QR of the stack, A mean
    return gemm(rows_out, nc, nc) + qr(nc + rows_out, rows_out) + gemm(rows_out, nc, mc)


def _revert(nc, mc, k):
    # (A chol)^T, QR of the (k + nc) block, gain solve, A mean, gain @ mean_observed
    return gemm(k, nc, nc) + qr(k + nc, k + nc) + trsm(k, nc) + gemm(k, nc, mc) + gemm(nc, k, mc)


def _merge(nc, mc):
    return gemm(nc, nc, nc) + gemm(nc, nc, mc) + gemm(nc, nc, nc) + qr(2 * nc, nc)


def attempt(cfg, n_tan: int = 0) -> float:
    """FLOPs of one attempted step (solver step + residual error estimate)."""
    nc, mc, ko, nb = dims(cfg, n_tan)
    f = 0.0
    lin = _marginalise(nc, mc, ko) + trsm(ko, mc)  # fx.marginalise(u) + whitened residual
    if cfg.constraint == "ts1":
        lin += gemm(ko, nc, 1)  # b = r(xi) - J xi
    if cfg.solver == "dynamic":
        f += gemm(nc, nc, mc)  # apply_flat (sigma = 1)
        f += lin
    if cfg.strategy == "filter":
        f += _marginalise(nc, mc, nc)
    else:
        f += _revert(nc, mc, nc)
        if cfg.strategy == "fixedpoint":
            f += _merge(nc, mc)
    f += _revert(nc, mc, ko)  # the update
    if cfg.solver != "dynamic":
        f += gemm(nc, nc, mc) + lin  # error estimate re-extrapolates and marginalises
    return nb * f


def interpolation(cfg) -> float:
    """FLOPs of one checkpoint interpolation beyond t1 (interpolate_fwd)."""
    nc, mc, _, nb = dims(cfg)
    if cfg.strategy == "filter":
        return nb * _marginalise(nc, mc, nc)
    f = 2 * _revert(nc, mc, nc)
    if cfg.strategy == "fixedpoint":
        f += _merge(nc, mc)
    return nb * f


def finalize_per_checkpoint(cfg) -> float:
    """Smoother backward pass: one conditional marginalisation per checkpoint."""
    nc, mc, _, nb = dims(cfg)
    if cfg.strategy == "filter":
        return 0.0
    return nb * _marginalise(nc, mc, nc)


def solve(cfg, attempts: float, checkpoints: int) -> float:
    """FLOPs of one solve with the given attempt count and number of checkpoints (K - 1)."""
    return attempts * attempt(cfg) + checkpoints * (interpolation(cfg) + finalize_per_checkpoint(cfg))


def lml_forward(cfg, grid_steps: int, observations: int, n_tan: int = 0) -> float:
    """Forward solve on the grid plus evaluate_lml's reverse scan; with n_tan > 0 the mean
    recursions carry the tangent columns of a forward-mode gradient (the factors are shared)."""
    nc, mc, ko, nb = dims(cfg, n_tan)
    step = attempt(cfg, n_tan) - (gemm(nc, nc, mc) + _marginalise(nc, mc, ko) + trsm(ko, mc))  # no error estimate
    lml = observations * (_marginalise(nc, mc, nc) + _revert(nc, mc, ko) + qr(ko, ko) + trsm(ko, mc))
    return nb * (grid_steps * step) + nb * lml


ADJ_SEGMENT = 50  # grid steps per checkpoint segment of the reverse adjoint


def lml_reverse(cfg, grid_steps: int, observations: int) -> float:
    """FLOPs the reverse-mode adjoint of the isotropic lane LML executes: the forward pass, the
    recomputation of every checkpoint segment and of every step's factors in the backward sweep
    (three forward passes' worth of factor work), and the adjoint of the mean recursions, which
    costs what one more set of mean columns costs in the forward direction (step and scan)."""
    nc, mc, ko, nb = dims(cfg)
    step0 = lml_forward(cfg, grid_steps, 0, 0) / max(1, grid_steps)
    step1 = lml_forward(cfg, grid_steps, 0, 1) / max(1, grid_steps)
    scan0 = lml_forward(cfg, 0, observations, 0)
    scan1 = lml_forward(cfg, 0, observations, 1)
    return nb * (3 * grid_steps * step0 + grid_steps * (step1 - step0) + scan0 + (scan1 - scan0))


def lml_reverse_bytes(cfg, n_inst: int, observations: int, grid_points: int) -> float:
    """Algorithmic HBM bytes of the reverse adjoint: inputs and outputs as in the forward LML
    (the Taylor coefficients carry their n_params + d sensitivities), and per instance its storage
    written once and read once: the interval conditionals, each observation node's gain, factor
    and residuals, the adjoints of the conditional means, the checkpoints every 50 steps, and
    every step's slot within its segment."""
    N, d = cfg.num_derivatives + 1, cfg.ode_dim
    tri = N * (N + 1) // 2
    slot = 8 + N * d + tri + N * N + N * d + tri + 2 * N
    cond = N * N + N * d + tri + 2 * N
    steps = grid_points - 1
    per_inst = (observations - 1) * cond + observations * (N + 1 + d) + (observations - 1) * N * d
    per_inst += ((steps - 1) // ADJ_SEGMENT + 1) * slot + steps * slot
    n_theta = cfg.n_params + d
    inputs = n_inst * (cfg.n_params + d + observations * (d + 1) + N * d * (1 + n_theta)) + grid_points
    outputs = n_inst * (1 + n_theta)
    return 8.0 * (inputs + outputs + 2 * n_inst * per_inst)


def solve_bytes(cfg, n_inst: int, K: int) -> float:
    """Algorithmic HBM bytes of one batched save_at solve: inputs, outputs written
    once, and the stored backward conditionals read back by the smoother's backward pass."""
    nc, mc, _, nb = dims(cfg)
    N, d = cfg.num_derivatives + 1, cfg.ode_dim
    ms = N * d                       # mean entries
    cs = nb * nc * nc                # factor entries
    inputs = cfg.n_params + N * d + K
    per_ckpt = 1 + ms + cs + nb + 1  # t, mean, chol, output_scale, num_steps
    reread = 0
    if cfg.strategy != "filter":
        cond = nb * nc * nc + ms + cs + 2 * nb * nc  # A, noise mean, noise chol, to_latent, to_observed
        per_ckpt += ms + cs + cond                   # filtering marginal + backward conditional
        reread = (K - 1) * cond
    return 8.0 * n_inst * (inputs + K * per_ckpt + reread)


def lml_bytes(cfg, n_inst: int, observations: int, grid_points: int, n_tan: int = 0) -> float:
    """Algorithmic HBM bytes of one batched LML (+ gradient): inputs (params, y0, the grid, data
    and one std per observation), outputs (value, gradient), and the backward conditional of every
    interval between observations (factors once, mean columns with their tangents) written by the
    forward pass and read back by the reverse scan."""
    nc, mc, _, nb = dims(cfg, n_tan)
    d = cfg.ode_dim
    cond = nb * (2 * nc * nc + 2 * nc) + nc * mc  # A, noise factor, to_latent, to_observed; noise mean
    inputs = n_inst * (cfg.n_params + d + observations * (d + 1)) + grid_points
    outputs = n_inst * (1 + n_tan)
    return 8.0 * (inputs + outputs + 2 * n_inst * (observations - 1) * cond)


def offgrid(cfg) -> float:
    """FLOPs of one off-grid marginal (solvers.py:149-203): the filter marginalises the transition;
    the fixed-interval smoother reverts it twice and marginalises one backward conditional."""
    nc, mc, _, nb = dims(cfg)
    if cfg.strategy == "filter":
        return nb * _marginalise(nc, mc, nc)
    return nb * (2 * _revert(nc, mc, nc) + _marginalise(nc, mc, nc))


def offgrid_bytes(cfg, n_query: int) -> float:
    """Algorithmic HBM bytes per batch of off-grid queries: each query reads its time, the two
    bracketing marginals (the filter needs one) and an output scale, and writes one marginal.  The
    binary search over the grid is not counted (it hits the cache)."""
    nc, mc, _, nb = dims(cfg)
    normal = nc * mc + nb * nc * nc
    reads = 1 + 1 + normal * (1 if cfg.strategy == "filter" else 2)
    return 8.0 * n_query * (reads + normal)
