"""The oracle's offgrid_marginals (solvers.py:149-203) against textbook Gaussian closed forms.

In observed (Taylor-coefficient) coordinates the IWP(nu) transition over dt with output scale s is
A(dt)_ij = dt^(j-i) / (j-i)!  (j >= i) and
Q(dt)_ij = s^2 |dt|^(2nu+1-i-j) / ((nu-i)! (nu-j)! (2nu+1-i-j))
(the reference's transition test, test_wiener_integrated.py:12-81).  The filter's off-grid marginal is
the prediction (A m0, A S0 A^T + Q); the fixed-interval smoother's is the Rauch-Tung-Striebel step
through t between the filtering marginal at t0 and the smoothed marginal at t1.  Covariances are
compared as L L^T (the factor's column signs are the QR's business).
"""

from math import factorial

import numpy as np
import pytest

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def _iwp(nu, dt, s):
    N = nu + 1
    A = np.zeros((N, N))
    Q = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            if j >= i:
                A[i, j] = dt ** (j - i) / factorial(j - i)
            p = 2 * nu + 1 - i - j
            Q[i, j] = s**2 * abs(dt) ** p / (factorial(nu - i) * factorial(nu - j) * p)
    return A, Q


def _random_normal(rng, n, cols):
    m = rng.normal(size=(n, cols))
    L = np.tril(rng.normal(size=(n, n))) + 2.0 * np.eye(n)
    return m, L


def _dense(nu, d, dt, s):
    A, Q = _iwp(nu, dt, s)
    return np.kron(A, np.eye(d)), np.kron(Q, np.eye(d))


@pytest.mark.parametrize("fact", ["isotropic", "dense"])
def test_filter_offgrid_is_the_gaussian_prediction(oracle, fact):
    rng = np.random.default_rng(1)
    nu, d, s = 3, 2, 0.7
    t0, t, t1 = 1.0, 1.37, 1.9
    cfg = oracle.Config(fact=fact, nu=nu, d=d, strategy="filter")
    if fact == "isotropic":
        m0, L0 = _random_normal(rng, nu + 1, d)
        A, Q = _iwp(nu, t - t0, s)
    else:
        m0, L0 = _random_normal(rng, (nu + 1) * d, 1)
        A, Q = _dense(nu, d, t - t0, s)
    mo, Lo = oracle.offgrid_marginal(cfg, THETA, t0, t1, t, s, m0.ravel(), L0.ravel())
    mo, Lo = np.asarray(mo).reshape(m0.shape), np.asarray(Lo).reshape(L0.shape)
    assert np.allclose(mo, A @ m0, rtol=1e-12, atol=1e-12)
    cov = A @ L0 @ L0.T @ A.T + Q
    assert np.allclose(Lo @ Lo.T, cov, rtol=1e-11, atol=1e-12 * np.abs(cov).max())


@pytest.mark.parametrize("fact", ["isotropic", "dense"])
def test_fixedinterval_offgrid_is_the_rts_step_through_t(oracle, fact):
    rng = np.random.default_rng(2)
    nu, d, s = 3, 2, 1.3
    t0, t, t1 = 0.5, 0.81, 1.2
    cfg = oracle.Config(fact=fact, nu=nu, d=d, strategy="fixedinterval")
    n = nu + 1 if fact == "isotropic" else (nu + 1) * d
    cols = d if fact == "isotropic" else 1
    m0, L0 = _random_normal(rng, n, cols)
    m1, L1 = _random_normal(rng, n, cols)
    if fact == "isotropic":
        A1, Q1 = _iwp(nu, t - t0, s)
        A2, Q2 = _iwp(nu, t1 - t, s)
    else:
        A1, Q1 = _dense(nu, d, t - t0, s)
        A2, Q2 = _dense(nu, d, t1 - t, s)
    mo, Lo = oracle.offgrid_marginal(cfg, THETA, t0, t1, t, s, m0.ravel(), L0.ravel(), m1.ravel(), L1.ravel())
    mo, Lo = np.asarray(mo).reshape(m0.shape), np.asarray(Lo).reshape(L0.shape)
    S0 = L0 @ L0.T
    mt, St = A1 @ m0, A1 @ S0 @ A1.T + Q1
    mp, Sp = A2 @ mt, A2 @ St @ A2.T + Q2
    G = np.linalg.solve(Sp, A2 @ St).T
    m_exp = mt + G @ (m1 - mp)
    S_exp = St + G @ (L1 @ L1.T - Sp) @ G.T
    assert np.allclose(mo, m_exp, rtol=1e-9, atol=1e-9 * np.abs(m_exp).max())
    assert np.allclose(Lo @ Lo.T, S_exp, rtol=1e-8, atol=1e-9 * np.abs(S_exp).max())


def test_offgrid_is_refused_for_the_fixedpoint_smoother(oracle):
    cfg = oracle.Config(fact="isotropic", nu=3, d=2, strategy="fixedpoint")
    m0, L0 = np.zeros((4, 2)), np.eye(4)
    with pytest.raises(NotImplementedError):
        oracle.offgrid_marginal(cfg, THETA, 0.0, 1.0, 0.5, 1.0, m0.ravel(), L0.ravel())


@pytest.mark.parametrize("fact,idx", [("isotropic", 0), ("isotropic", 2), ("dense", 0), ("dense", 1), ("blockdiag", 0)])
def test_lml_terminal_values_is_the_gaussian_log_density(oracle, fact, idx):
    """loss_lml_terminal_values (estimators_and_losses.py:20-50): the log-density of u under
    derivative idx of the marginal plus independent noise, against scipy."""
    from scipy.stats import multivariate_normal, norm

    rng = np.random.default_rng(5 + idx)
    nu, d = 3, 2
    N = nu + 1
    cfg = oracle.Config(fact=fact, nu=nu, d=d, strategy="filter")
    u = rng.normal(size=d)
    if fact == "isotropic":
        m, L = _random_normal(rng, N, d)
        std = 0.3
        var = L[idx] @ L[idx] + std**2
        expected = sum(norm.logpdf(u[k], m[idx, k], np.sqrt(var)) for k in range(d))
        got = oracle.lml_terminal(cfg, THETA, m.ravel(), L.ravel(), u, std, idx)
    elif fact == "dense":
        m, L = _random_normal(rng, N * d, 1)
        std = np.array([0.3, 0.5])
        rows = L[idx * d : (idx + 1) * d]
        cov = rows @ rows.T + np.diag(std**2)
        expected = multivariate_normal(m[idx * d : (idx + 1) * d, 0], cov).logpdf(u)
        got = oracle.lml_terminal(cfg, THETA, m.ravel(), L.ravel(), u, std, idx)
    else:
        ms = [rng.normal(size=N) for _ in range(d)]
        Ls = [np.tril(rng.normal(size=(N, N))) + 2 * np.eye(N) for _ in range(d)]
        std = np.array([0.3, 0.5])
        expected = sum(norm.logpdf(u[k], ms[k][idx], np.sqrt(Ls[k][idx] @ Ls[k][idx] + std[k] ** 2)) for k in range(d))
        got = oracle.lml_terminal(cfg, THETA, np.concatenate(ms), np.concatenate([x.ravel() for x in Ls]), u, std, idx)
    assert np.isclose(got, expected, rtol=1e-13, atol=1e-13)
