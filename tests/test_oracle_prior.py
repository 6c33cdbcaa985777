"""The oracle's exponential prior (ssm_impl_dense.py:392-629, util/gram_util.py) against closed forms.

* exp_gram_cholesky(pade_and_legendre_9) against the matrix-fraction baseline of
  test_util.exp_gram_matrix_fraction (test_gram_util.py:34-71): e^A and the Gramian
  G = int_0^1 e^{sA} B B^T e^{sA^T} ds are the blocks E[:n, :n] and E[:n, n:] e^{A^T} of
  E = expm([[A, B B^T], [0, -A^T]]) (scipy's expm), for the reference's small-norm inputs and for
  norms that need several doublings.  The tolerances are a few times the measured errors, so a
  wrong Pade coefficient, a wrong Legendre normalisation or a missing doubling fails; the highest
  Legendre terms are below FP64 resolution (a 10 % change of C[2, 8] moves the Gramian by 8e-20
  relative in the long double build) and no FP64 result depends on them.
* With zero drift the exponential prior is the integrated Wiener process
  (test_exponential.py:8-30): the same transition at dt = +-1.234 and output scale 12.3456.
* Proposition 1 of arXiv:2305.14978 (test_exponential.py:58-92): in the original coordinates the
  transition's last block is expm(M dt) and its leading block is the IWP transition of one
  derivative fewer.
* Whole solves: the zero-drift exponential prior reproduces the IWP solve (same steps), and a
  Matern prior solves Lotka-Volterra to the tolerance (DOP853 reference).
"""

from math import comb

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def _matrix_fraction(A, B):
    n = A.shape[0]
    M = np.block([[A, B @ B.T], [np.zeros_like(A), -A.T]])
    E = expm(M)
    eA = E[:n, :n]
    return eA, E[:n, n:] @ eA.T


@pytest.mark.parametrize("ncols", [2, 8])
@pytest.mark.parametrize("norm", [0.01, "eta", 3.0, 40.0])
@pytest.mark.parametrize("triu", [True, False])
def test_exp_gram_matches_matrix_fraction(oracle, ncols, norm, triu):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(8, 8)) / np.sqrt(8)
    if triu:
        A = np.triu(A)
    # "eta": 1-norm 0.4, just below the degree-9 threshold 0.41: no doubling, and the
    # approximant's highest terms weigh the most
    A = A * (0.4 / np.max(np.sum(np.abs(A), axis=0)) if norm == "eta" else norm)
    B = rng.normal(size=(8, ncols))
    eA, U = oracle.exp_gram(A, B)
    eA1, G1 = _matrix_fraction(A, B)
    # measured: 1e-15 (norm 0.01), 1e-14 (norm 3), 1e-13 (norm 40, 7e-12 where the baseline's own
    # Gramian is the less accurate of the two); the doublings square e^A, so errors grow with the norm
    tol = {0.01: 2e-14, "eta": 2e-14, 3.0: 1e-13, 40.0: 2e-11}[norm]
    assert np.max(np.abs(eA - eA1)) <= tol * np.max(np.abs(eA1))
    assert np.max(np.abs(U @ U.T - G1)) <= tol * np.max(np.abs(G1))
    assert np.array_equal(U, np.tril(U))
    assert np.all(np.diag(U) >= 0)


def test_exp_gram_of_a_scalar_ou_process(oracle):
    """e^{-l} and the OU variance (1 - e^{-2 l}) / (2 l) of dX = -l X dt + dW over unit time."""
    for lam in [0.3, 2.0, 25.0]:
        eA, U = oracle.exp_gram(np.array([[-lam]]), np.array([[1.0]]))
        assert abs(eA[0, 0] - np.exp(-lam)) <= 1e-14 * np.exp(-lam) + 1e-300
        var = (1 - np.exp(-2 * lam)) / (2 * lam)
        assert abs(U[0, 0] ** 2 - var) <= 1e-13 * var


def _exp_cfg(oracle, nu, d, drift, **kw):
    return oracle.Config(fact="dense", nu=nu, d=d, prior="exponential", prior_drift=drift, **kw)


@pytest.mark.parametrize("dt", [1.234, -1.234])
def test_exponential_prior_with_zero_drift_matches_iwp(oracle, dt):
    nu, d = 3, 2
    n = (nu + 1) * d
    scale = 12.3456
    A1, L1, tl1, to1 = oracle.transition(_exp_cfg(oracle, nu, d, np.zeros((d, n))), dt, scale)
    A2, L2, tl2, to2 = oracle.transition(oracle.Config(fact="dense", nu=nu, d=d), dt, scale)
    assert np.array_equal(tl1, tl2) and np.array_equal(to1, to2)
    assert np.allclose(A1, A2, rtol=0, atol=1e-12 * np.max(np.abs(A2)))
    C1, C2 = L1 @ L1.T, L2 @ L2.T
    assert np.max(np.abs(C1 - C2)) <= 1e-12 * np.max(np.abs(C2))


@pytest.mark.parametrize("dt", [1.234, -1.234])
@pytest.mark.parametrize("d", [1, 3])
def test_exponential_transition_contains_expm(oracle, dt, d):
    nu = 2
    n = (nu + 1) * d
    rng = np.random.default_rng(1)
    M = rng.normal(size=(d, d))
    drift = np.zeros((d, n))
    drift[:, nu * d :] = M  # f(u, du, ddu) = M ddu
    A, _, tl, to = oracle.transition(_exp_cfg(oracle, nu, d, drift), dt, 1.0)
    A_orig = to[:, None] * A * tl[None, :]  # preconditioner_apply
    assert np.allclose(A_orig[-d:, -d:], expm(M * dt), rtol=1e-10, atol=1e-12)
    Ai, _, tli, toi = oracle.transition(oracle.Config(fact="dense", nu=nu - 1, d=d), dt, 1.0)
    assert np.allclose(A_orig[:-d, :-d], toi[:, None] * Ai * tli[None, :], rtol=1e-10, atol=1e-12)


def test_zero_drift_solve_reproduces_the_iwp_solve(oracle):
    nu, d = 4, 2
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, nu)
    save_at = np.linspace(0.0, 20.0, 50)
    r_iwp = oracle.solve_save_at(oracle.Config(fact="dense", nu=nu), THETA, tc, save_at)
    r_exp = oracle.solve_save_at(_exp_cfg(oracle, nu, d, np.zeros((d, (nu + 1) * d))), THETA, tc, save_at)
    assert r_iwp["attempts"] == r_exp["attempts"]
    assert np.array_equal(r_iwp["num_steps"], r_exp["num_steps"])
    assert np.max(np.abs(r_exp["mean"] - r_iwp["mean"])) <= 1e-8 * np.max(np.abs(r_iwp["mean"]))


def matern_drift(length_scale, N, d):
    """prior_matern (ssm_impl_api.py:443-477): f(u_0..u_{N-1}) = -sum_i C(N, i) z^(N-i) u_i."""
    z = np.sqrt(2 * (N - 0.5)) / length_scale
    return np.concatenate([-comb(N, i) * z ** (N - i) * np.eye(d) for i in range(N)], axis=1)


def test_matern_drift_matches_the_reference_closed_form():
    """test_matern.py:7-27: for three coefficients, f = -l^3 u - 3 l^2 du - 3 l ddu, l = sqrt(5) / ls."""
    ls = 0.567
    ell = np.sqrt(2 * 2.5) / ls
    expected = np.concatenate([-(ell**3) * np.eye(4), -3 * ell**2 * np.eye(4), -3 * ell * np.eye(4)], axis=1)
    assert np.allclose(matern_drift(ls, 3, 4), expected, rtol=1e-15)


def test_matern_prior_solves_lotka_volterra(oracle):
    nu, d = 4, 2
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, nu)
    save_at = np.linspace(0.0, 20.0, 50)
    r = oracle.solve_save_at(_exp_cfg(oracle, nu, d, matern_drift(10.0, nu + 1, d), rtol=1e-6, atol=1e-8), THETA, tc, save_at)
    assert r["status"] == 0
    ref = solve_ivp(lambda t, y: [THETA[0] * y[0] - THETA[1] * y[0] * y[1], -THETA[2] * y[1] + THETA[3] * y[0] * y[1]],
                    (0, 20), [20.0, 20.0], method="DOP853", rtol=1e-13, atol=1e-13, t_eval=save_at).y.T
    u = r["mean"].reshape(50, -1)[:, :2]
    assert np.max(np.abs(u - ref) / np.abs(ref)) < 1e-4
    r_iwp = oracle.solve_save_at(oracle.Config(fact="dense", nu=nu), THETA, tc, save_at)
    assert r["attempts"] != r_iwp["attempts"] or not np.allclose(r["mean"], r_iwp["mean"], rtol=1e-12)

