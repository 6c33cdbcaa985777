"""Shared inputs and checkers for the marginal-log-likelihood tests.

The gradient is pinned by Richardson-extrapolated central differences of the
oracle's value, (4 D(h) - D(2h)) / 3 with D(h) = (f(x+h) - f(x-h)) / 2h and
h = 1e-4 |x|, evaluated with the long double build of the oracle: truncation
O(h^4) and rounding both stay near 1e-13 relative.  On these inputs the kernel
code (emulated on the CPU) agrees with it to 1e-12 normwise or better; the
tolerance is 1e-10.  (With FP64 function values the step must be ~10x larger
to keep rounding down, and the truncation error then reaches ~1e-8.)
"""

import numpy as np

THETA = np.array([0.5, 0.05, 0.5, 0.05])
GRAD_RTOL = 1e-10


def lml_inputs(n_inst, seed=0, t_end=20.0, n_grid=2001, every=20, per_coordinate_std=False):
    """Noisy Lotka-Volterra observations at grid[::every]; parameters and y0 perturbed off the truth."""
    from scipy.integrate import solve_ivp

    rng = np.random.default_rng(seed)
    p_true = THETA * np.exp(rng.uniform(np.log(0.8), np.log(1.2), size=(n_inst, 4)))
    y_true = rng.uniform(10.0, 30.0, size=(n_inst, 2))
    grid = np.linspace(0.0, t_end, n_grid)
    t_obs = grid[::every]
    data = np.zeros((n_inst, t_obs.size, 2))
    for i in range(n_inst):
        a, b, c, d = p_true[i]
        s = solve_ivp(lambda t, y: [a * y[0] - b * y[0] * y[1], -c * y[1] + d * y[0] * y[1]], (0, t_end), y_true[i], method="DOP853", rtol=1e-10, atol=1e-10, t_eval=t_obs)
        data[i] = s.y.T
    data = data + 0.1 * np.abs(data) * rng.normal(size=data.shape)
    if per_coordinate_std:
        stds = 0.1 * np.abs(data) + 0.5
    else:
        stds = (0.1 * np.sqrt(np.mean(data**2, axis=-1)))[..., None]
    params = p_true * np.exp(rng.uniform(-0.1, 0.1, size=p_true.shape))
    y0 = y_true + rng.uniform(-1.0, 1.0, size=y_true.shape)
    return params, y0, grid, every, data, stds


def lml_inputs_scale_sensitive(n_inst, seed=0):
    """A coarse grid (dt = 0.25) and observation stds of 1e-3: the solver's own uncertainty
    dominates the observed covariances, so the LML depends on the posterior's output scale."""
    params, y0, grid, every, data, _ = lml_inputs(n_inst, seed=seed, t_end=6.0, n_grid=25, every=2)
    return params, y0, grid, every, data, np.full(data.shape[:2] + (1,), 1e-3)


def oracle_value(oracle, ocfg, params, y0, grid, every, data, stds, average=True):
    """(FP64 strict, FP64 FMA, long double) values of the oracle's LML on the same float64 inputs."""
    v = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    with oracle.rounding("fma"):
        vf = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    with oracle.rounding("extended"):
        vt = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    return v, vf, vt


def richardson_grad(oracle, ocfg, params, y0, grid, every, data, stds, average=True, step=1e-4):
    """Richardson-extrapolated central differences (steps h and 2h) of the oracle's LML, evaluated
    in the long double build: its values carry about 19 digits, so a step of 1e-4 |x| keeps both the
    rounding error and the O(h^4) truncation error far below the 1e-10 tolerance the kernels are held to.

    For solver_dynamic the output scales are recorded at x and replayed at every perturbed point,
    which differentiates with the calibration held constant: the reference's stop_gradient
    (solvers.py:568-569), the gradient the kernels compute."""
    x = np.concatenate([params, y0]).astype(np.longdouble)
    n_p = params.size
    dynamic = ocfg.solver == "dynamic"

    with oracle.rounding("extended"):
        tape = np.zeros((np.asarray(grid).size - 1) * max(1, y0.size), dtype=np.longdouble)
        if dynamic:
            oracle.lml_from_y0_tape(ocfg, x[:n_p], x[n_p:], grid, every, data, stds, tape, True, average)

    def f(z):
        if dynamic:
            return np.longdouble(oracle.lml_from_y0_tape(ocfg, z[:n_p], z[n_p:], grid, every, data, stds, tape, False, average))
        return np.longdouble(oracle.lml_from_y0(ocfg, z[:n_p], z[n_p:], grid, every, data, stds, average))

    g = np.zeros(x.size, dtype=np.longdouble)
    with oracle.rounding("extended"):
        for j in range(x.size):
            h = np.longdouble(step) * max(abs(x[j]), np.longdouble(1e-3))
            e = np.zeros(x.size, dtype=np.longdouble)
            e[j] = h
            d1 = (f(x + e) - f(x - e)) / (2 * h)
            d2 = (f(x + 2 * e) - f(x - 2 * e)) / (4 * h)
            g[j] = (4 * d1 - d2) / 3
    return g.astype(np.float64)


def grad_reference(oracle, ocfg, params, y0, grid, every, data, stds, average=True):
    """(Richardson gradient at step 1e-4, its tolerance): the tolerance is GRAD_RTOL or, where the
    oracle's LML is less well conditioned (data observing a high Taylor coefficient), ten times the
    spread between the steps 1e-4 and 1e-3, the finite differences' own measured uncertainty."""
    g = richardson_grad(oracle, ocfg, params, y0, grid, every, data, stds, average)
    g3 = richardson_grad(oracle, ocfg, params, y0, grid, every, data, stds, average, step=1e-3)
    return g, max(GRAD_RTOL, 10.0 * rel(g, g3))


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def lml_dropped_singular_value_inputs(n_inst=2, seed=3):
    """Dense LML data on a coarse grid (11 points on [0, 5], every point observed) where coordinate 1
    of every third observation has std 1e17 (how missing data are often encoded) and coordinate 0
    there has std 0.1.  The observed factor R_Y at those nodes has singular values ~1e17 and ~1:
    lstsq_svd's cutoff eps * 2 * s_max (~44) drops the small one, so its gain ignores coordinate 0
    as well, while solve_triu's gain uses it.  The grid is coarse so that the state is uncertain
    enough at the nodes for the gain to matter to the likelihood."""
    params, y0, grid, every, data, stds = lml_inputs(n_inst, seed=seed, t_end=5.0, n_grid=11, every=1, per_coordinate_std=True)
    stds[:, 1::3, 1] = 1e17
    stds[:, 1::3, 0] = 0.1
    return params, y0, grid, every, data, stds
