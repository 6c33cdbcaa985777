"""Shared cases for the off-grid marginal tests (solvers.py:149-203).

A case is one Lotka-Volterra solve by the oracle (FP64, strict order): the filter on a save_at grid,
the fixed-interval smoother on a fixed grid.  Query times lie strictly inside the grid intervals and
off its points, as the reference requires.  The oracle's off-grid marginal of every query in its
three builds (FP64 strict, FP64 with FMA, long double) is the reference the kernels are held to.
"""

import numpy as np

import test_gpu_parity as T


def make_case(oracle, nu, strategy, solver="dynamic", n_query=12, seed=0, fact="isotropic", **cfg_kw):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation=fact, num_derivatives=nu, ode_dim=2, strategy=strategy, solver=solver, **cfg_kw)
    ocfg = T._oracle_cfg(oracle, cfg)
    params, y0, _ = T._lv_batch(1, seed=seed)
    tc = oracle.taylor_coefficients("lotka_volterra", params[0], y0[0], 0.0, nu)
    if strategy == "filter":
        times = np.linspace(0.0, 6.0, 13)
        sol = oracle.solve_save_at(ocfg, params[0], tc, times)
    else:
        times = np.linspace(0.0, 3.0, 61)
        sol = oracle.solve_fixed_grid(ocfg, params[0], tc, times)
    K = times.size
    os_ = np.array(sol["output_scale"], dtype=np.float64).reshape(-1)
    if os_.size == K - 1:  # non-dynamic solvers: K - 1 rows in the reference; index K - 1 clamps to the last
        os_ = np.append(os_, os_[-1])
    os_ = os_.reshape(K, 1)
    rng = np.random.default_rng(seed + 100)
    idx = rng.integers(1, K, size=n_query)
    frac = rng.uniform(0.05, 0.95, size=n_query)
    tq = times[idx - 1] + frac * (times[idx] - times[idx - 1])
    solution = {"mean": np.asarray(sol["mean"], dtype=np.float64), "chol": np.asarray(sol["chol"], dtype=np.float64), "output_scale": os_}
    if strategy != "filter":
        solution["filt_mean"] = np.asarray(sol["filt_mean"], dtype=np.float64)
        solution["filt_chol"] = np.asarray(sol["filt_chol"], dtype=np.float64)
    refs = []
    for q in range(n_query):
        i = int(idx[q])
        m0 = solution["mean" if strategy == "filter" else "filt_mean"][i - 1]
        L0 = solution["chol" if strategy == "filter" else "filt_chol"][i - 1]
        m1, L1 = solution["mean"][i], solution["chol"][i]
        per = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                m, L = oracle.offgrid_marginal(ocfg, params[0], times[i - 1], times[i], tq[q], os_[i, 0], m0, L0, m1, L1)
            per.append((np.asarray(m, dtype=np.float64), np.asarray(L, dtype=np.float64)))
        refs.append(per)
    return cfg, times[None], solution, tq[None], refs


def check(got_mean, got_chol, refs, label, cov=False):
    """cov: compare covariances L L^T instead of the factors (the dense factors' column signs can
    hang on cancellation-level entries)."""

    def c(L):
        L = np.asarray(L, dtype=np.float64)
        n = int(round(np.sqrt(L.size)))
        L = L.reshape(n, n)
        return L @ L.T if cov else L

    for q, per in enumerate(refs):
        (mr, Lr), (mf, Lf), (mt, Lt) = per
        T.check_accuracy(got_mean[q], mr, mf, mt, f"{label} query {q} mean")
        T.check_accuracy(c(got_chol[q]), c(Lr), c(Lf), c(Lt), f"{label} query {q} {'cov' if cov else 'chol'}")
