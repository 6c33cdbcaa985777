"""Pin the oracle's Monte Carlo Jacobian handlers (jacobians.py:116-271).

The reference draws Rademacher probes from JAX's threefry stream, which is not
available here; the oracle and the kernels share a counter-based stream instead
(one probe set per linearisation, indexed by the accepted steps).  What the
reference fixes, and these tests check:

* the estimator: calculate_diagonal_along_d is the probe mean of vjp(v) * v
  (reverse mode, :250-271) or jvp(v) * v (forward mode, :171-192), recomputed
  here with numpy matrix products on the oracle's own probes;
* for a diagonal Jacobian both estimates are exact (v_k^2 = 1);
* the estimate is unbiased: over many probe sets it tends to the Jacobian's
  diagonal, and its sum over k to the trace;
* the probes are balanced signs, independent across coordinates;
* the key travels with the accepted solution (a rejected proposal's key is
  discarded), so an adaptive solve replayed on its own accepted grid by
  solve_fixed_grid reproduces it (test_solve_fixed_grid.py:17-54 with the
  handler switched on);
* for a diagonal Jacobian a whole ts1 solve with either handler equals the
  exact-Jacobian solve.
"""

import numpy as np
import pytest

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def _probes(oracle, seed, set_, probes, d):
    return np.array([[oracle.rademacher(seed, set_, s, j) for j in range(d)] for s in range(probes)])


def test_probe_stream_is_balanced_and_independent(oracle):
    v = np.array([[oracle.rademacher(7, st, s, j) for j in range(8)] for st in range(40) for s in range(16)])
    assert set(np.unique(v)) == {-1.0, 1.0}
    n = v.shape[0]
    # means and pairwise correlations of independent fair signs are O(1 / sqrt(n))
    assert np.max(np.abs(v.mean(axis=0))) < 5.0 / np.sqrt(n)
    corr = (v.T @ v) / n - np.eye(8)
    assert np.max(np.abs(corr)) < 5.0 / np.sqrt(n)
    # different seeds and probe sets give different probes
    a, b, c = _probes(oracle, 1, 0, 10, 6), _probes(oracle, 2, 0, 10, 6), _probes(oracle, 1, 1, 10, 6)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


CASES = [
    ("lotka_volterra", THETA, np.array([20.0, 15.0])),
    ("van_der_pol", np.array([3.0]), np.array([1.5, -0.7])),
    ("burgers", np.array([0.02, 0.25]), np.array([0.3, -0.1, 0.8, 0.5])),
    ("burgers", np.array([0.02, 1.0 / 7.0]), np.sin(np.linspace(0.0, 2.0 * np.pi, 7, endpoint=False)) + 0.2),
    ("burgers", np.array([0.05, 1.0 / 3.0]), np.array([0.4, -0.3, 0.9])),
]


@pytest.mark.parametrize("mode", ["rev", "fwd"])
@pytest.mark.parametrize("vf,prm,y", CASES, ids=[f"{c[0]}-{c[2].size}" for c in CASES])
def test_diagonal_estimate_equals_vjp_jvp_products(oracle, vf, prm, y, mode):
    d = y.size
    _, J = oracle.vector_field(vf, prm, y)
    assert not np.allclose(J, J.T)  # a transposed product would be caught
    for seed, set_, probes in [(1, 0, 10), (3, 17, 5), (1, 4, 64)]:
        V = _probes(oracle, seed, set_, probes, d)
        prod = V @ J if mode == "rev" else V @ J.T  # rows: v^T J (vjp) or (J v)^T (jvp)
        want = np.mean(prod * V, axis=0)
        got = oracle.mc_jacobian_diagonal(vf, prm, y, seed=seed, set_=set_, probes=probes, mode=mode)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.max(np.abs(J)))


@pytest.mark.parametrize("mode", ["rev", "fwd"])
def test_diagonal_jacobian_is_estimated_exactly(oracle, mode):
    for vf, prm, y in [("linear", np.array([-0.5, 0.3, -2.0]), np.array([1.0, 2.0, 3.0])),
                       ("logistic", np.array([2.0]), np.array([0.1, 0.4, 0.9, 0.5]))]:
        _, J = oracle.vector_field(vf, prm, y)
        got = oracle.mc_jacobian_diagonal(vf, prm, y, seed=5, set_=3, probes=10, mode=mode)
        np.testing.assert_allclose(got, np.diag(J), rtol=1e-15, atol=0.0)


@pytest.mark.parametrize("mode", ["rev", "fwd"])
def test_estimate_is_unbiased(oracle, mode):
    vf, prm, y = CASES[3]
    d = y.size
    _, J = oracle.vector_field(vf, prm, y)
    sets = 600
    est = np.array([oracle.mc_jacobian_diagonal(vf, prm, y, seed=1, set_=s, probes=10, mode=mode) for s in range(sets)])
    # per-entry variance of one probe: sum_{j != k} J[j][k]^2 (rev) or J[k][j]^2 (fwd)
    off = J - np.diag(np.diag(J))
    var = np.sum(off**2, axis=0 if mode == "rev" else 1) / 10.0
    se = np.sqrt(var / sets)
    assert np.all(np.abs(est.mean(axis=0) - np.diag(J)) < 5.0 * se)
    # single estimates scatter (the estimator is live) around the trace
    tr = est.sum(axis=1)
    assert np.std(tr) > 0.0
    assert abs(tr.mean() - np.trace(J)) < 5.0 * np.std(tr) / np.sqrt(sets)


@pytest.mark.parametrize("fact", ["isotropic", "blockdiag"])
@pytest.mark.parametrize("mode", ["monte_carlo_rev", "monte_carlo_fwd"])
@pytest.mark.parametrize("re_lin", [False, True])
def test_adaptive_solve_replays_on_its_grid(oracle, fact, mode, re_lin):
    """Rejected attempts keep the accepted solution's probe set: the fixed-grid replay of the
    accepted steps draws the same probes and reproduces the adaptive solve."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 3)
    cfg = oracle.Config(fact=fact, nu=3, strategy="filter", solver="dynamic", constraint="ts1", rtol=1e-4,
                        atol=1e-6, clip_dt=True, jacobian=mode, jac_probes=4, jac_seed=11,
                        re_linearize_after_calibration=re_lin)
    ad = oracle.solve_save_every_step(cfg, THETA, tc, 0.0, 4.0)
    assert ad["status"] == 0 and ad["attempts"] > ad["accepted"]  # rejections happened
    # the replay's steps t_k - t_{k-1} round differently from the adaptive dt: 1e-10 of each
    # output's scale, as in test_fixed_grid_reproduces_adaptive_grid (four probes make a noisy
    # linearisation, whose highest derivatives grow large)
    fx = oracle.solve_fixed_grid(cfg, THETA, tc, ad["t"])
    for key in ("mean", "chol"):
        assert np.max(np.abs(fx[key] - ad[key])) <= 1e-10 * max(1.0, np.max(np.abs(ad[key])))
    # and the probes matter: the exact trace, or another seed's probes, take other paths
    for change in ({"jacobian": "exact"}, {"jac_seed": 12}):
        other = oracle.solve_fixed_grid(oracle.Config(**{**cfg.__dict__, **change}), THETA, tc, ad["t"])
        assert np.max(np.abs(other["mean"] - ad["mean"])) > 1e-6 * max(1.0, np.max(np.abs(ad["mean"])))


@pytest.mark.parametrize("fact", ["isotropic", "blockdiag"])
@pytest.mark.parametrize("mode", ["monte_carlo_rev", "monte_carlo_fwd"])
def test_diagonal_ode_solve_equals_exact_jacobian(oracle, fact, mode):
    lam = np.array([-0.5, -1.5, 0.25])
    tc = oracle.taylor_coefficients("linear", lam, [1.0, 0.5, -1.0], 0.0, 3)
    save_at = np.linspace(0.0, 2.0, 9)
    base = dict(fact=fact, nu=3, strategy="fixedpoint", solver="dynamic", constraint="ts1", d=3, vf="linear",
                rtol=1e-6, atol=1e-8)
    ex = oracle.solve_save_at(oracle.Config(**base), lam, tc, save_at)
    mc = oracle.solve_save_at(oracle.Config(**base, jacobian=mode, jac_probes=7), lam, tc, save_at)
    assert ex["status"] == mc["status"] == 0
    assert ex["attempts"] == mc["attempts"]
    np.testing.assert_allclose(mc["mean"], ex["mean"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mc["chol"], ex["chol"], rtol=1e-12, atol=1e-12)
