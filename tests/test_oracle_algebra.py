"""Pin the oracle's square-root algebra, IWP matrices and small ops.

Each check compares the oracle against something independent of it: LAPACK
(numpy) for QR signs, Kalman closed forms, mpmath exact Cholesky factors,
the reference's own closed-form KATs, scipy densities and sympy derivatives.
"""

import mpmath
import numpy as np
import pytest
import scipy.stats
import sympy


def _rng(seed=1):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("shape", [(20, 10), (12, 2), (10, 10), (6, 6), (24, 12), (5, 1), (8, 4), (3, 5)])
def test_qr_r_matches_lapack_including_signs(oracle, shape):
    """backend/linalg.py:8-37: qr_r is LAPACK geqrf's R (numpy uses the same routine)."""
    M = _rng().normal(size=shape)
    R = oracle.qr_r(M)
    R_np = np.linalg.qr(M, mode="r")
    assert R.shape == R_np.shape
    assert np.allclose(R, R_np, rtol=1e-12, atol=1e-12)


def test_qr_r_zero_subcolumns_leave_the_column_untouched(oracle):
    """dlarfg with a zero sub-column: tau = 0, R[i, i] = alpha (sign kept)."""
    M = np.zeros((6, 3))
    M[0, 0] = -2.0
    M[1:, 1] = 0.0
    M[1, 2] = 3.0
    R = oracle.qr_r(M)
    assert R[0, 0] == -2.0
    assert np.all(R[:, 1] == 0.0)
    assert np.allclose(R.T @ R, M.T @ M)
    R_np = np.linalg.qr(M, mode="r")
    assert np.allclose(R, R_np)


@pytest.mark.parametrize(("HCshape", "Cshape", "Xshape"), [((4, 3), (3, 3), (4, 4)), ((2, 3), (3, 3), (2, 2)), ((1, 5), (5, 5), (1, 1))])
@pytest.mark.parametrize("lstsq", [False, True])
def test_revert_conditional_kalman(oracle, HCshape, Cshape, Xshape, lstsq):
    """tests/test_util/test_cholesky_util.py:14-33 restated: S, K and the backward noise."""
    rng = _rng()
    HC = rng.normal(size=HCshape) + 1.0
    C = rng.normal(size=Cshape) + 2.0
    X = rng.normal(size=Xshape) + 3.0 + np.eye(Xshape[0])
    S = HC @ HC.T + X @ X.T
    K = C @ HC.T @ np.linalg.inv(S)
    C1 = C @ C.T - K @ S @ K.T
    extra, (bw_noise, g) = oracle.revert_conditional(HC.T, C.T, X.T, lstsq=lstsq)
    assert np.allclose(extra.T @ extra, S, rtol=1e-10)
    assert np.allclose(g, K, rtol=1e-10, atol=1e-12)
    assert np.allclose(bw_noise.T @ bw_noise, C1, rtol=1e-9, atol=1e-10)


def test_lstsq_gain_is_minimum_norm_for_singular_factor(oracle):
    """lstsq_svd semantics: a zero observation factor yields the pseudo-inverse (zero) gain."""
    n, k = 3, 1
    R_X_F = np.zeros((n, k))
    R_X = np.zeros((n, n))
    R_YX = np.zeros((k, k))
    R_Y, (R_XY, G) = oracle.revert_conditional(R_X_F, R_X, R_YX, lstsq=True)
    assert np.all(G == 0.0)
    assert np.all(np.isfinite(R_XY))


@pytest.mark.parametrize("nu", [1, 2, 3, 4, 5, 7])
def test_iwp_system_matrices(oracle, nu):
    """utilities.py:57-71: A = flip(Pascal), Q = lower Cholesky of the flipped Hilbert matrix."""
    A, Q = oracle.system_matrices(nu)
    N = nu + 1
    A_expected = np.array([[float(sympy.binomial(nu - i, nu - j)) if j >= i else 0.0 for j in range(N)] for i in range(N)])
    assert np.array_equal(A, A_expected)
    mpmath.mp.dps = 50
    H = mpmath.matrix(N, N)
    for i in range(N):
        for j in range(N):
            H[i, j] = mpmath.mpf(1) / (2 * nu + 1 - i - j)
    L = mpmath.cholesky(H)
    L_expected = np.array([[float(L[i, j]) for j in range(N)] for i in range(N)])
    assert np.allclose(Q, L_expected, rtol=1e-12, atol=1e-15)
    assert np.all(np.diag(Q) > 0)


@pytest.mark.parametrize("fact", ["dense", "isotropic", "blockdiag"])
@pytest.mark.parametrize("dt", [1.234, -1.234])
def test_transitions_are_correct_in_1d(oracle, fact, dt):
    """tests/test_probdiffeq/test_priors/test_wiener_integrated.py:12-81 (nu=3, d=1)."""
    cfg = oracle.Config(fact=fact, nu=3, d=1)
    A, L, tl, to = oracle.transition(cfg, dt)
    A_eff = to[:, None] * A * tl[None, :]
    A_expected = np.asarray(
        [
            [1.0, dt, dt**2 / 2.0, dt**3 / 6.0],
            [0.0, 1.0, dt, dt**2 / 2.0],
            [0.0, 0.0, 1.0, dt],
            [0.0, 0.0, 0.0, 1.0],
        ]
    )
    assert np.allclose(A_eff, A_expected, rtol=1e-13)
    h = abs(dt)
    Q_expected = np.asarray(
        [
            [h**7 / 252.0, h**6 / 72.0, h**5 / 30.0, h**4 / 24.0],
            [h**6 / 72.0, h**5 / 20.0, h**4 / 8.0, h**3 / 6.0],
            [h**5 / 30.0, h**4 / 8.0, h**3 / 3.0, h**2 / 2.0],
            [h**4 / 24.0, h**3 / 6.0, h**2 / 2.0, h**1 / 1.0],
        ]
    )
    Lp = np.abs(to)[:, None] * L
    assert np.allclose(Lp @ Lp.T, Q_expected, rtol=1e-12)


@pytest.mark.parametrize("fact", ["dense", "isotropic", "blockdiag"])
def test_output_scales_scale_the_transition_covariance(oracle, fact):
    """tests/test_probdiffeq/test_priors/test_output_scales.py:93-122."""
    cfg = oracle.Config(fact=fact, nu=1, d=1, base_scale=123.45)
    _, L, _, to = oracle.transition(cfg, 1.0, os=9.876)
    Lp = np.abs(to)[:, None] * L
    Q_expected = (9.876 * 123.45) ** 2.0 * 1.0 / np.asarray([[3.0, 2.0], [2.0, 1.0]])
    assert np.allclose(Lp @ Lp.T, Q_expected, rtol=1e-12)


def test_dense_transition_is_kron_of_1d(oracle):
    """ssm_impl_dense.py:505-508: A = kron(a, I_d), Q = kron(Q1, I_d), derivative-major."""
    cfg = oracle.Config(fact="dense", nu=2, d=3)
    A, L, tl, to = oracle.transition(cfg, 0.7)
    cfg1 = oracle.Config(fact="isotropic", nu=2, d=3)
    A1, L1, tl1, to1 = oracle.transition(cfg1, 0.7)
    assert np.allclose(A, np.kron(A1, np.eye(3)))
    assert np.allclose(L, np.kron(L1, np.eye(3)))
    assert np.allclose(tl, np.repeat(tl1, 3)) and np.allclose(to, np.repeat(to1, 3))


@pytest.mark.parametrize("isotropic", [False, True])
def test_logpdf_matches_scipy(oracle, isotropic):
    """tests/test_probdiffeq/test_logpdf.py:18-29."""
    rng = _rng(3)
    n, d = 5, 3
    L = np.tril(rng.normal(size=(n, n))) + 3 * np.eye(n)
    if isotropic:
        mean = rng.normal(size=(n, d))
        u = np.ones((n, d))
        expected = scipy.stats.multivariate_normal.logpdf(u.T.reshape(-1), mean.T.reshape(-1), np.kron(np.eye(d), L @ L.T))
    else:
        mean = rng.normal(size=n)
        u = np.ones(n)
        expected = scipy.stats.multivariate_normal.logpdf(u, mean, L @ L.T)
    got = oracle.logpdf(mean, L, u, isotropic=isotropic)
    assert np.isclose(got, expected, rtol=1e-12)


def test_controller_pi_with_zero_proportional_gain_is_integral(oracle):
    """tests/test_ivpsolve/test_controllers.py:10-26 (dt=0.1428, error_power=3.142, 4 applies)."""
    c_pi = oracle.Config(control="proportional_integral", exponent_integral=1.0, exponent_proportional=0.0)
    c_i = oracle.Config(control="integral")
    dt_pi, x_pi, dt_i = 0.1428, 1.0, 0.1428
    for _ in range(4):
        dt_pi, x_pi = oracle.control(c_pi, dt_pi, x_pi, 3.142)
        dt_i, _ = oracle.control(c_i, dt_i, 0.0, 3.142)
    assert np.isclose(dt_i, dt_pi, rtol=1e-15)
    # safety 0.95 * 3.142 = 2.9849 per apply, below the clip at 10
    assert np.isclose(dt_i, 0.1428 * (0.95 * 3.142) ** 4, rtol=1e-14)


def test_controller_clips(oracle):
    c = oracle.Config()
    assert oracle.control(c, 1.0, 0.0, 1e6)[0] == 10.0
    assert oracle.control(c, 1.0, 0.0, 1e-6)[0] == 0.2


def _sympy_derivatives(f_expr, ys, y0, num, t=None, t0=0.0):
    """Derivatives d^k y / dt^k at t0 for y' = f(y, t) by repeated total differentiation."""
    derivs = [sympy.Matrix(ys)]
    cur = sympy.Matrix(f_expr)
    derivs.append(cur)
    for _ in range(num - 1):
        nxt = cur.jacobian(ys) * sympy.Matrix(f_expr)
        if t is not None:
            nxt = nxt + cur.diff(t)
        derivs.append(nxt)
        cur = nxt
    subs = {y: v for y, v in zip(ys, y0)}
    if t is not None:
        subs[t] = t0
    return np.array([[float(e.subs(subs)) for e in D] for D in derivs[: num + 1]])


def test_taylor_coefficients_lotka_volterra(oracle):
    a, b, c, d = 0.5, 0.05, 0.5, 0.05
    y1, y2 = sympy.symbols("y1 y2")
    f = [a * y1 - b * y1 * y2, -c * y2 + d * y1 * y2]
    expected = _sympy_derivatives(f, [y1, y2], [20.0, 20.0], 5)
    got = oracle.taylor_coefficients("lotka_volterra", [a, b, c, d], [20.0, 20.0], 0.0, 5)
    assert np.allclose(got, expected, rtol=1e-13)


def test_taylor_coefficients_van_der_pol(oracle):
    mu = 7.5
    y1, y2 = sympy.symbols("y1 y2")
    f = [y2, mu * ((1 - y1**2) * y2) - y1]
    expected = _sympy_derivatives(f, [y1, y2], [2.0, 0.3], 6)
    got = oracle.taylor_coefficients("van_der_pol", [mu], [2.0, 0.3], 0.0, 6)
    assert np.allclose(got, expected, rtol=1e-12)


def test_taylor_coefficients_time_dependent_linear(oracle):
    lam = [0.3, -1.2]
    y1, y2, t = sympy.symbols("y1 y2 t")
    f = [lam[0] * y1 + t, lam[1] * y2 + t]
    expected = _sympy_derivatives(f, [y1, y2], [1.0, -2.0], 4, t=t, t0=0.5)
    got = oracle.taylor_coefficients("linear", lam, [1.0, -2.0], 0.5, 4)
    assert np.allclose(got, expected, rtol=1e-13)


def test_taylor_coefficients_burgers_small_grid(oracle):
    N = 6
    dx = 1.0 / N
    visc = 0.01
    us = sympy.symbols(f"u0:{N}")
    f = [-us[j] * ((us[(j + 1) % N] - us[(j - 1) % N]) / (2 * dx)) + visc * ((us[(j + 1) % N] - 2 * us[j] + us[(j - 1) % N]) / dx**2) for j in range(N)]
    u0 = list(np.sin(2 * np.pi * np.arange(N) / N))
    expected = _sympy_derivatives(f, list(us), u0, 3)
    got = oracle.taylor_coefficients("burgers", [visc, dx], u0, 0.0, 3)
    assert np.allclose(got, expected, rtol=1e-11, atol=1e-9)


def test_vector_field_jacobians_match_finite_differences(oracle):
    for vf, prm, y in [
        ("lotka_volterra", [0.5, 0.05, 0.5, 0.05], [20.0, 13.0]),
        ("van_der_pol", [30.0], [1.5, -0.4]),
        ("logistic", [2.0], [0.3]),
        ("burgers", [0.01, 1 / 8], list(np.cos(np.arange(8)))),
    ]:
        f, J = oracle.vector_field(vf, prm, y)
        h = 1e-6
        Jfd = np.zeros_like(J)
        for j in range(len(y)):
            yp, ym = np.array(y, float), np.array(y, float)
            yp[j] += h
            ym[j] -= h
            Jfd[:, j] = (oracle.vector_field(vf, prm, yp)[0] - oracle.vector_field(vf, prm, ym)[0]) / (2 * h)
        assert np.allclose(J, Jfd, rtol=1e-6, atol=1e-6), vf


@pytest.mark.parametrize(
    "case",
    ["full_rank", "rank_one_triangular", "zero", "cutoff_hit", "cutoff_missed", "tall_rank_deficient", "rank_two_of_three"],
)
def test_lstsq_svd_matches_numpy(oracle, case):
    """lstsq_svd (backend/linalg.py:60-61) is jnp.linalg.lstsq(A, B)[0]: the minimum-norm solution
    with singular values at or below eps * max(m, n) * s_max dropped, which numpy.linalg.lstsq with
    its default rcond computes by LAPACK's SVD.  The cutoff cases sit a factor 10 either side of the
    threshold (eps * 2 * 1e16 ~ 4.4): a wrong cutoff (relative to s_min, or without max(m, n)) or a
    back substitution instead fails one of them."""
    rng = _rng()
    if case == "full_rank":
        A = np.triu(rng.normal(size=(3, 3))) + 2.0 * np.eye(3)
    elif case == "rank_one_triangular":
        A = np.array([[1.0, 2.0], [0.0, 0.0]])
    elif case == "zero":
        A = np.zeros((2, 2))
    elif case == "cutoff_hit":
        A = np.array([[1e16, 3.0], [0.0, 0.44]])
    elif case == "cutoff_missed":
        A = np.array([[1e16, 3.0], [0.0, 44.0]])
    elif case == "tall_rank_deficient":
        u = rng.normal(size=(5, 1))
        A = np.hstack([u, 2.0 * u, rng.normal(size=(5, 1))])
    else:
        A = rng.normal(size=(3, 2)) @ rng.normal(size=(2, 3))
    B = rng.normal(size=(A.shape[0], 4))
    X = oracle.lstsq_svd(A, B)
    X_np = np.linalg.lstsq(A, B, rcond=None)[0]
    assert np.allclose(X, X_np, rtol=1e-12, atol=1e-12 * np.abs(X_np).max()), (X, X_np)
    if case == "cutoff_hit":
        # the small singular value is dropped: the solution has no component along it
        assert not np.allclose(X, np.linalg.solve(A, B), rtol=1e-3)
    if case == "cutoff_missed":
        assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-10)
