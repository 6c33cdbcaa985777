"""Shared inputs and checkers for the marginal-log-likelihood tests.

The gradient is pinned by Richardson-extrapolated central differences of the
oracle's value, (4 D(h) - D(2h)) / 3 with D(h) = (f(x+h) - f(x-h)) / 2h and
h = 1e-3 |x|: truncation O(h^4), rounding ~1e-13 relative.  On these inputs
the kernel code (emulated on the CPU) agrees with it to ~1e-10 normwise.
"""

import numpy as np

THETA = np.array([0.5, 0.05, 0.5, 0.05])
GRAD_RTOL = 1e-8


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


def oracle_value(oracle, ocfg, params, y0, grid, every, data, stds, average=True):
    """(FP64 strict, FP64 FMA, long double) values of the oracle's LML on the same float64 inputs."""
    v = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    with oracle.rounding("fma"):
        vf = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    with oracle.rounding("extended"):
        vt = oracle.lml_from_y0(ocfg, params, y0, grid, every, data, stds, average)
    return v, vf, vt


def richardson_grad(oracle, ocfg, params, y0, grid, every, data, stds, average=True):
    x = np.concatenate([params, y0])
    n_p = params.size

    def f(z):
        return oracle.lml_from_y0(ocfg, z[:n_p], z[n_p:], grid, every, data, stds, average)

    g = np.zeros(x.size)
    for j in range(x.size):
        h = 1e-3 * max(abs(x[j]), 1e-3)
        e = np.zeros(x.size)
        e[j] = h
        d1 = (f(x + e) - f(x - e)) / (2 * h)
        d2 = (f(x + 2 * e) - f(x - 2 * e)) / (4 * h)
        g[j] = (4 * d1 - d2) / 3
    return g


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
