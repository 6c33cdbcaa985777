"""The oracle's error estimators (solvers.py:850-1098) on one step, against closed forms.

Isotropic IWP(nu) prior, ts0 for Lotka-Volterra, solver_dynamic, filter.  With output scale 1 the
extrapolation of the previous mean m is rv = N(A m, Q) (A(dt), Q(dt) the IWP transition in
Taylor-coefficient coordinates, test_wiener_integrated.py:12-81).  The step's linearisation (the
same point: the calibration extrapolates the mean with output scale 1) observes derivative 1:
residual z = (A m)_1 - f((A m)_0) with variance S = Q_11 (+ damp^2), output scale
sigma = ||z|| / sqrt(S d).
* error_residual_std: error = sigma sqrt(S), order n = 1;
* error_state_std(derivative_idx=i): error = sigma sqrt((Q - q q^T / S)_ii) with q = Q[:, 1],
  order n = i;
error_per_unit_step adds one to n; error_abs = error dt^n / n!; the error power is
rms(error_abs / (atol + rtol max(|m_i|, |m+_i|)))^(-1/(nu+1)) with m+ the proposed mean.
"""

from math import factorial

import numpy as np
import pytest

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def _iwp(nu, dt):
    N = nu + 1
    A = np.zeros((N, N))
    Q = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            if j >= i:
                A[i, j] = dt ** (j - i) / factorial(j - i)
            p = 2 * nu + 1 - i - j
            Q[i, j] = abs(dt) ** p / (factorial(nu - i) * factorial(nu - j) * p)
    return A, Q


def _lv(y, th):
    a, b, c, d = th
    return np.array([a * y[0] - b * y[0] * y[1], -c * y[1] + d * y[0] * y[1]])


@pytest.mark.parametrize(
    "estimator,idx,per_unit,relin",
    [("residual_std", 0, False, False), ("residual_std", 0, True, True), ("state_std", 0, False, False), ("state_std", 3, False, True), ("state_std", 2, True, False)],
)
def test_error_power_of_one_step_matches_the_closed_form(oracle, estimator, idx, per_unit, relin):
    nu, d = 3, 2
    N = nu + 1
    rng = np.random.default_rng(idx + 3 * per_unit)
    m = np.zeros((N, d))
    m[0] = [20.0, 12.0]
    m[1] = _lv(m[0], THETA)
    m[2:] = rng.normal(size=(N - 2, d))
    L = np.tril(rng.normal(size=(N, N))) * 1e-3
    dt, atol, rtol = 0.07, 1e-6, 1e-4
    cfg = oracle.Config(fact="isotropic", nu=nu, d=d, strategy="filter", solver="dynamic", atol=atol, rtol=rtol,
                        error_estimator=estimator, error_derivative_idx=idx, error_per_unit_step=per_unit, error_re_linearize=relin)
    ep, prop = oracle.step_error_power(cfg, THETA, m.ravel(), L.ravel(), 0.0, dt)
    prop = np.asarray(prop).reshape(N, d)
    A, Q = _iwp(nu, dt)
    mr = A @ m
    z = mr[1] - _lv(mr[0], THETA)
    S = Q[1, 1]
    sigma = np.sqrt(np.sum(z**2) / S / d)
    if estimator == "residual_std":
        err, n, row = sigma * np.sqrt(S), 1, 0
    else:
        q = Q[:, 1]
        post = Q - np.outer(q, q) / S
        err, n, row = sigma * np.sqrt(post[idx, idx]), idx, idx
    if per_unit:
        n += 1
    err_abs = err * dt**n / factorial(n)
    ref = np.maximum(np.abs(m[row]), np.abs(prop[row]))
    norm = np.sqrt(np.mean((err_abs / (atol + rtol * ref)) ** 2))
    assert np.isclose(float(ep), norm ** (-1.0 / N), rtol=1e-9)


@pytest.mark.parametrize("relin", [False, True])
def test_state_std_solves_are_accurate(oracle, relin):
    """As the reference's solve test matrix does (test_solve_adaptive_save_at.py:174-185): a solve with
    error_state_std at rtol 1e-6 meets the tolerance against DOP853 (rtol 1e-13)."""
    from scipy.integrate import solve_ivp

    cfg = oracle.Config(fact="isotropic", nu=4, d=2, strategy="fixedpoint", solver="dynamic", rtol=1e-6, atol=1e-8,
                        error_estimator="state_std", error_re_linearize=relin)
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    save_at = np.linspace(0.0, 20.0, 50)
    r = oracle.solve_save_at(cfg, THETA, tc, save_at)
    a, b, c, d = THETA
    ref = solve_ivp(lambda t, y: [a * y[0] - b * y[0] * y[1], -c * y[1] + d * y[0] * y[1]], (0, 20), [20.0, 20.0],
                    method="DOP853", rtol=1e-13, atol=1e-13, t_eval=save_at).y.T
    u = np.asarray(r["mean"]).reshape(50, 5, 2)[:, 0]
    assert np.max(np.abs(u - ref) / (1e-8 + 1e-6 * np.abs(ref))) < 50.0
