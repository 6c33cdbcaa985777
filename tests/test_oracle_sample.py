"""The oracle's MarkovSequence.sample (estimators_and_losses.py:233-273), pinned without statistics.

A sample is affine in its standard normal draws z: x = x(0) + M z.  For the dense factorisation
(apply_flat scales the whole mean by to_observed, ssm_impl_dense.py:18-22) x(0) must be the
smoothing means and M M^T, assembled from unit draws, the smoothing covariances that the solve's
backward pass computes.  The isotropic sampler is the dense one on rewritten inputs: its apply_flat
adds the noise mean without to_observed (ssm_impl_isotropic.py:75-79) and its sample_flat draws one
N-vector for all d columns (:255-258), which the dense sampler reproduces with to_observed folded
into A and each draw repeated d times.
"""

import numpy as np
import pytest

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def _solve(oracle, fact, K=7, nu=3):
    cfg = oracle.Config(fact=fact, nu=nu, d=2, strategy="fixedpoint", solver="dynamic")
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, nu)
    sol = oracle.solve_save_at(cfg, THETA, tc, np.linspace(0.0, 3.0, K))
    return cfg, sol


def test_dense_samples_have_the_smoothing_marginals(oracle):
    cfg, sol = _solve(oracle, "dense")
    K, n = np.asarray(sol["mean"]).shape
    x0 = np.asarray(oracle.sample(cfg, THETA, sol, np.zeros((K, n))))
    means = np.asarray(sol["mean"])
    assert np.max(np.abs(x0 - means)) <= 1e-11 * np.max(np.abs(means))
    cols = []
    for j in range(K * n):
        z = np.zeros(K * n)
        z[j] = 1.0
        cols.append(np.asarray(oracle.sample(cfg, THETA, sol, z.reshape(K, n))) - x0)
    M = np.stack(cols, axis=-1)  # (K, n, K n)
    for k in range(K):
        L = np.asarray(sol["chol"][k]).reshape(n, n)
        cov = L @ L.T
        assert np.allclose(M[k] @ M[k].T, cov, rtol=1e-8, atol=1e-10 * np.max(np.abs(cov))), k


def test_isotropic_sampler_is_the_dense_one_on_the_rewritten_conditionals(oracle):
    cfg_i, sol = _solve(oracle, "isotropic")
    N, d = cfg_i.nu + 1, cfg_i.d
    K = np.asarray(sol["mean"]).shape[0]
    rng = np.random.default_rng(0)
    z = rng.normal(size=(K, N))
    xi = np.asarray(oracle.sample(cfg_i, THETA, sol, z)).reshape(K, N, d)
    I = np.eye(d)
    A = np.asarray(sol["cond_A"]).reshape(K - 1, N, N)
    to = np.asarray(sol["cond_to"]).reshape(K - 1, N)
    tl = np.asarray(sol["cond_tl"]).reshape(K - 1, N)
    nL = np.asarray(sol["cond_chol"]).reshape(K - 1, N, N)
    dense = {
        "term_mean": np.asarray(sol["term_mean"]).reshape(N * d),
        "term_chol": np.kron(np.asarray(sol["term_chol"]).reshape(N, N), I),
        "cond_A": np.stack([np.kron(to[k][:, None] * A[k], I) for k in range(K - 1)]),
        "cond_mean": np.asarray(sol["cond_mean"]).reshape(K - 1, N * d),
        "cond_chol": np.stack([np.kron(np.abs(to[k])[:, None] * nL[k], I) for k in range(K - 1)]),
        "cond_tl": np.repeat(tl, d, axis=1),
        "cond_to": np.ones((K - 1, N * d)),
    }
    cfg_d = oracle.Config(fact="dense", nu=cfg_i.nu, d=d, strategy="fixedpoint", solver="dynamic")
    xd = np.asarray(oracle.sample(cfg_d, THETA, dense, np.repeat(z, d, axis=1))).reshape(K, N, d)
    assert np.allclose(xi, xd, rtol=1e-13, atol=1e-13 * np.max(np.abs(xi)))


def test_sampling_refuses_blockdiag(oracle):
    cfg = oracle.Config(fact="blockdiag", nu=2, d=2, strategy="fixedpoint")
    sol = {k: np.zeros(4) for k in ("term_mean", "term_chol", "cond_A", "cond_mean", "cond_chol", "cond_tl", "cond_to")}
    with pytest.raises(NotImplementedError):
        oracle.sample(cfg, THETA, sol, np.zeros((2, 3)))
