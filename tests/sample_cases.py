"""Shared cases for the posterior-sampling tests (estimators_and_losses.py:233-273): an oracle
fixed-point solve, standard normal draws, and the oracle's samples in its three builds."""

import numpy as np

import test_gpu_parity as T


def make_case(oracle, nu, n_samples=6, K=20, seed=0, fact="isotropic", strategy="fixedpoint", **cfg_kw):
    """base: (n_samples, K, n) draws, n the covariance dimension ((nu + 1) d dense, nu + 1 isotropic)."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation=fact, num_derivatives=nu, ode_dim=2, strategy=strategy, solver="dynamic", **cfg_kw)
    ocfg = T._oracle_cfg(oracle, cfg)
    params, y0, _ = T._lv_batch(1, seed=seed)
    tc = oracle.taylor_coefficients("lotka_volterra", params[0], y0[0], 0.0, nu)
    times = np.linspace(0.0, 5.0, K)
    sol = oracle.solve_save_at(ocfg, params[0], tc, times)
    keys = ("term_mean", "term_chol", "cond_A", "cond_mean", "cond_chol", "cond_tl", "cond_to")
    sol = {k: np.asarray(sol[k], dtype=np.float64) for k in keys}
    rng = np.random.default_rng(seed + 7)
    base = rng.normal(size=(n_samples, K, cfg.n))
    refs = []
    for s in range(n_samples):
        per = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                per.append(np.asarray(oracle.sample(ocfg, params[0], sol, base[s]), dtype=np.float64))
        refs.append(per)
    return cfg, sol, base, refs


def check(samples, refs, label):
    for s, (r, f, t) in enumerate(refs):
        T.check_accuracy(samples[s], r, f, t, f"{label} sample {s}")
