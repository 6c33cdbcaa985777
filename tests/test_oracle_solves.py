"""Pin the oracle's whole solves.

Against: scipy's high-accuracy integrators (DOP853 / Radau at rtol <= 1e-10),
convergence with the tolerance, the reference's own equivalence tests
(dense == isotropic, vmap(dense) == blockdiag, fixed-point == fixed-interval,
fixed grid == adaptive grid, truncated save_at == prefix) and a brute-force
joint-Gaussian evaluation of the marginal log-likelihood.
"""

import numpy as np
import pytest
from scipy.integrate import solve_ivp

THETA = np.array([0.5, 0.05, 0.5, 0.05])


def lv(t, y, p=THETA):
    a, b, c, d = p
    return [a * y[0] - b * y[0] * y[1], -c * y[1] + d * y[0] * y[1]]


def lv_reference(save_at, y0=(20.0, 20.0), p=THETA):
    s = solve_ivp(lv, (save_at[0], save_at[-1]), y0, method="DOP853", rtol=1e-13, atol=1e-13, t_eval=save_at, args=(p,))
    return s.y.T


@pytest.mark.parametrize("fact", ["dense", "isotropic", "blockdiag"])
@pytest.mark.parametrize("strategy", ["filter", "fixedpoint"])
def test_lotka_volterra_matches_dop853(oracle, fact, strategy):
    """test_solve_adaptive_save_at.py:202-247 at the 50-time-unit horizon (C1 tolerances)."""
    cfg = oracle.Config(fact=fact, strategy=strategy, nu=4, rtol=1e-6, atol=1e-8)
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    save_at = np.linspace(0.0, 50.0, 200)
    r = oracle.solve_save_at(cfg, THETA, tc, save_at)
    assert r["status"] == 0
    mean0 = r["mean"].reshape(200, -1)
    if fact == "dense":
        u = mean0[:, :2]
    elif fact == "isotropic":
        u = mean0.reshape(200, 5, 2)[:, 0, :]
    else:
        u = mean0.reshape(200, 2, 5)[:, :, 0]
    ref = lv_reference(save_at)
    assert np.max(np.abs(u - ref) / np.abs(ref)) < 1e-4
    assert np.allclose(r["t"], save_at, rtol=0, atol=1e-12)
    assert r["num_steps"][-1] == r["accepted"]


def test_error_decreases_with_tolerance(oracle):
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    save_at = np.linspace(0.0, 10.0, 11)
    ref = lv_reference(save_at)
    errs = []
    for tol in [1e-4, 1e-6, 1e-8]:
        cfg = oracle.Config(rtol=tol, atol=1e-2 * tol)
        r = oracle.solve_save_at(cfg, THETA, tc, save_at)
        errs.append(np.max(np.abs(r["mean"][:, :2] - ref)))
    assert errs[0] > errs[1] > errs[2]


def test_van_der_pol_ts1_matches_radau(oracle):
    mu = 10.0
    cfg = oracle.Config(fact="dense", nu=5, constraint="ts1", vf="van_der_pol", rtol=1e-5, atol=1e-7)
    tc = oracle.taylor_coefficients("van_der_pol", [mu], [2.0, 0.0], 0.0, 5)
    r = oracle.solve_save_at(cfg, [mu], tc, np.array([0.0, 6.3]))
    s = solve_ivp(lambda t, y: [y[1], mu * ((1 - y[0] ** 2) * y[1]) - y[0]], (0, 6.3), [2.0, 0.0], method="Radau", rtol=1e-12, atol=1e-12)
    assert np.allclose(r["mean"][-1][:2], s.y[:, -1], rtol=1e-4, atol=1e-5)


def test_burgers_blockdiag_matches_dop853(oracle):
    N = 16
    x = np.arange(N) / N
    u0 = np.sin(2 * np.pi * x) + 0.5 * np.sin(4 * np.pi * x)
    prm = [0.01, 1.0 / N]
    cfg = oracle.Config(fact="blockdiag", nu=2, d=N, vf="burgers", rtol=1e-6, atol=1e-8)
    tc = oracle.taylor_coefficients("burgers", prm, u0, 0.0, 2)
    r = oracle.solve_save_at(cfg, prm, tc, np.array([0.0, 0.5]))

    def fb(t, u):
        up, um = np.roll(u, -1), np.roll(u, 1)
        return -u * ((up - um) / (2 * prm[1])) + prm[0] * ((up - 2 * u + um) / prm[1] ** 2)

    s = solve_ivp(fb, (0, 0.5), u0, method="DOP853", rtol=1e-12, atol=1e-12)
    u = r["mean"][-1].reshape(N, 3)[:, 0]
    assert np.max(np.abs(u - s.y[:, -1])) < 1e-4


@pytest.mark.parametrize("constraint,estimator", [("ts0", "residual_std"), ("ts1", "residual_std"), ("ts1", "state_std")])
def test_dense_equals_isotropic_for_isotropic_ode(oracle, constraint, estimator):
    """test_dynamic_across_factorisations.py:12-65: for f(u) = lam u + t the dense ts1 Jacobian row
    [-lam I, I, 0] is kron([-tr(lam I)/d, 1, 0], I), so the isotropic ts1 linearisation (exact trace,
    jacobian_materialize) and the dense one describe the same model.  The reference's test observes
    the jet-lifted ODE (all derivatives); this one the plain first-order residual."""
    lam = [-0.7] * 5
    u0 = np.linspace(-1, 1, 5)
    save_at = np.linspace(0.0, 10.0, 7)
    res = []
    for fact in ["dense", "isotropic"]:
        cfg = oracle.Config(fact=fact, nu=2, d=5, vf="linear", strategy="fixedpoint", rtol=1e-3, atol=1e-3,
                            constraint=constraint, error_estimator=estimator, error_derivative_idx=1)
        tc = oracle.taylor_coefficients("linear", lam, u0, 0.0, 2)
        res.append(oracle.solve_save_at(cfg, lam, tc, save_at))
    rd, ri = res
    assert rd["accepted"] == ri["accepted"]
    assert np.allclose(rd["output_scale"], ri["output_scale"], rtol=1e-10)
    assert np.allclose(rd["mean"], ri["mean"], rtol=1e-10, atol=1e-12)
    # dense covariance == kron(iso covariance, I_d)
    for k in range(7):
        Ld = rd["chol"][k].reshape(15, 15)
        Li = ri["chol"][k].reshape(3, 3)
        assert np.allclose(Ld @ Ld.T, np.kron(Li @ Li.T, np.eye(5)), rtol=1e-9, atol=1e-12)


def test_isotropic_ts1_lotka_volterra_matches_dop853(oracle):
    """Isotropic ts1 (exact trace) meets the tolerance like ts0 (test_solve_adaptive_save_at.py:202-247),
    and differs from ts0 (the Jacobian term is live)."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    save_at = np.linspace(0.0, 50.0, 200)
    out = {}
    for constraint in ["ts0", "ts1"]:
        cfg = oracle.Config(fact="isotropic", strategy="fixedpoint", nu=4, rtol=1e-6, atol=1e-8, constraint=constraint)
        out[constraint] = r = oracle.solve_save_at(cfg, THETA, tc, save_at)
        assert r["status"] == 0
        u = r["mean"].reshape(200, 5, 2)[:, 0, :]
        ref = lv_reference(save_at)
        assert np.max(np.abs(u - ref) / np.abs(ref)) < 1e-4
    assert out["ts0"]["attempts"] != out["ts1"]["attempts"] or not np.array_equal(out["ts0"]["mean"], out["ts1"]["mean"])


def test_vmap_dense_equals_blockdiag_for_diagonal_ode(oracle):
    """test_dynamic_across_factorisations.py:68-145 (fixed grid, fixed-interval, ts0)."""
    rng = np.random.default_rng(1)
    diag = rng.normal(size=5)
    u0 = rng.normal(size=5)
    ts = np.linspace(0.0, 10.0, 10)
    cfg_bd = oracle.Config(fact="blockdiag", nu=3, d=5, vf="linear", strategy="fixedinterval")
    tc = oracle.taylor_coefficients("linear", diag, u0, 0.0, 3)
    rb = oracle.solve_fixed_grid(cfg_bd, diag, tc, ts)
    for j in range(5):
        cfg_d = oracle.Config(fact="dense", nu=3, d=1, vf="linear", strategy="fixedinterval")
        tcj = oracle.taylor_coefficients("linear", diag[j : j + 1], u0[j : j + 1], 0.0, 3)
        rd = oracle.solve_fixed_grid(cfg_d, diag[j : j + 1], tcj, ts)
        assert np.allclose(rd["mean"], rb["mean"].reshape(10, 5, 4)[:, j, :], rtol=1e-10, atol=1e-12)
        assert np.allclose(rd["chol"].reshape(10, 4, 4), rb["chol"].reshape(10, 5, 4, 4)[:, j], rtol=1e-10, atol=1e-12)
        assert np.allclose(rd["output_scale"][:, 0], rb["output_scale"][:, j], rtol=1e-12)


@pytest.mark.parametrize("vf", ["linear", "logistic"])
@pytest.mark.parametrize("strategy", ["filter", "fixedpoint"])
def test_vmap_dense_equals_blockdiag_ts1_for_diagonal_ode(oracle, vf, strategy):
    """Blockdiag ts1 (BlockDiagResidual.linearize, ssm_impl_blockdiag.py:147-183, with the exact
    Jacobian diagonal of jacobian_materialize, jacobians.py:108-113) against dense ts1 on each
    component: for an ODE whose Jacobian is diagonal the blocks are d independent dense d = 1
    problems that share the grid (test_dynamic_across_factorisations.py:68-145 with ts1).  The
    logistic field's Jacobian depends on the state, so a wrong sign or a missing diagonal entry
    in the block operator (or in its offset fx) changes the calibration and the means."""
    rng = np.random.default_rng(4)
    d = 5
    if vf == "linear":
        prm = rng.normal(size=d)
        u0 = rng.normal(size=d)
    else:
        prm = np.array([1.7])
        u0 = rng.uniform(0.1, 0.9, size=d)
    ts = np.linspace(0.0, 3.0, 13)
    cfg_bd = oracle.Config(fact="blockdiag", nu=3, d=d, vf=vf, constraint="ts1", strategy=strategy)
    tc = oracle.taylor_coefficients(vf, prm, u0, 0.0, 3)
    rb = oracle.solve_fixed_grid(cfg_bd, prm, tc, ts)
    cfg_b0 = oracle.Config(fact="blockdiag", nu=3, d=d, vf=vf, constraint="ts0", strategy=strategy)
    r0 = oracle.solve_fixed_grid(cfg_b0, prm, tc, ts)
    assert np.max(np.abs(r0["output_scale"] - rb["output_scale"])) > 1e-6  # ts1 is not ts0
    for j in range(d):
        pj = prm[j : j + 1] if vf == "linear" else prm
        cfg_d = oracle.Config(fact="dense", nu=3, d=1, vf=vf, constraint="ts1", strategy=strategy)
        tcj = oracle.taylor_coefficients(vf, pj, u0[j : j + 1], 0.0, 3)
        rd = oracle.solve_fixed_grid(cfg_d, pj, tcj, ts)
        assert np.allclose(rd["mean"], rb["mean"].reshape(13, d, 4)[:, j, :], rtol=1e-10, atol=1e-12)
        assert np.allclose(rd["chol"].reshape(13, 4, 4), rb["chol"].reshape(13, d, 4, 4)[:, j], rtol=1e-10, atol=1e-12)
        assert np.allclose(rd["output_scale"][:, 0], rb["output_scale"][:, j], rtol=1e-12)


def _marginalise_np(A, nm, nc, tl, to, m, L):
    """Dense LatentCond.marginalise (ssm_impl_dense.py:24-33) on covariances."""
    mean = tl[:, None] * m
    chol = tl[:, None] * L
    mean_new = to[:, None] * (A @ mean + nm)
    cov = (A @ chol) @ (A @ chol).T + nc @ nc.T
    cov = np.abs(to)[:, None] * cov * np.abs(to)[None, :]
    return mean_new, cov


@pytest.mark.parametrize("fact", ["dense", "isotropic"])
def test_fixedpoint_equals_fixedinterval_same_grid(oracle, fact):
    """test_smoother_fixedinterval_vs_fixedpoint.py:50-86 with the save-every-step harness."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 2)
    cfg_fi = oracle.Config(fact=fact, nu=2, strategy="fixedinterval", solver="solver", rtol=1e-3, atol=1e-3)
    sm = oracle.solve_save_every_step(cfg_fi, THETA, tc, 0.0, 2.0)
    cfg_fp = oracle.Config(fact=fact, nu=2, strategy="fixedpoint", solver="solver", rtol=1e-3, atol=1e-3)
    fp = oracle.solve_save_at(cfg_fp, THETA, tc, sm["t"])
    assert np.allclose(fp["t"], sm["t"])
    assert np.allclose(fp["mean"], sm["mean"], rtol=1e-9, atol=1e-9)
    assert np.array_equal(fp["num_steps"], sm["num_steps"])
    n = 6 if fact == "dense" else 3
    for k in range(len(sm["t"])):
        Ca = fp["chol"][k].reshape(n, n)
        Cb = sm["chol"][k].reshape(n, n)
        assert np.allclose(Ca @ Ca.T, Cb @ Cb.T, rtol=1e-8, atol=1e-12)


def test_smoother_marginals_are_an_rts_pass_over_the_filter(oracle):
    """MarkovSequence.evaluate_marginals (estimators_and_losses.py:156-178) recomputed in numpy."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 3)
    cfg = oracle.Config(fact="dense", nu=3, strategy="fixedpoint", solver="dynamic", rtol=1e-4, atol=1e-6)
    save_at = np.linspace(0, 3, 9)
    r = oracle.solve_save_at(cfg, THETA, tc, save_at)
    n = 8
    m = r["term_mean"].reshape(n, 1)
    L = r["term_chol"].reshape(n, n)
    cov = L @ L.T
    for j in range(len(save_at) - 2, -1, -1):
        A = r["cond_A"][j].reshape(n, n)
        nm = r["cond_mean"][j].reshape(n, 1)
        nc = r["cond_chol"][j].reshape(n, n)
        tl, to = r["cond_tl"][j], r["cond_to"][j]
        mean = tl[:, None] * m
        Ccov = (tl[:, None] * cov) * tl[None, :]
        m = to[:, None] * (A @ mean + nm)
        cov = np.abs(to)[:, None] * (A @ Ccov @ A.T + nc @ nc.T) * np.abs(to)[None, :]
        Lk = r["chol"][j].reshape(n, n)
        assert np.allclose(m[:, 0], r["mean"][j], rtol=1e-9, atol=1e-9)
        assert np.allclose(cov, Lk @ Lk.T, rtol=1e-6, atol=1e-10)


def test_fixed_grid_reproduces_adaptive_grid(oracle):
    """test_solve_fixed_grid.py:17-54: solve_fixed_grid on the adaptive grid (filter, MLE, clip_dt)."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 2)
    for fact in ["dense", "isotropic", "blockdiag"]:
        cfg = oracle.Config(fact=fact, nu=2, strategy="filter", solver="mle", rtol=1e-2, atol=1e-2, clip_dt=True)
        ad = oracle.solve_save_every_step(cfg, THETA, tc, 0.0, 2.0)
        fx = oracle.solve_fixed_grid(cfg, THETA, tc, ad["t"])
        assert np.allclose(ad["mean"], fx["mean"], rtol=1e-10, atol=1e-10)
        assert np.allclose(ad["chol"], fx["chol"], rtol=1e-10, atol=1e-10)
        assert np.allclose(ad["output_scale"], fx["output_scale"], rtol=1e-10)


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint"])
def test_truncated_save_at_is_a_prefix(oracle, strategy):
    """test_behaviour_close_to_t1.py:59-94 / truncated save_at == prefix."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    cfg = oracle.Config(strategy=strategy)
    save_at = np.linspace(0, 10, 21)
    full = oracle.solve_save_at(cfg, THETA, tc, save_at)
    part = oracle.solve_save_at(cfg, THETA, tc, save_at[:12])
    if strategy == "filter":
        assert np.array_equal(full["mean"][:12], part["mean"])
        assert np.array_equal(full["num_steps"][:11], part["num_steps"])
    else:
        # Smoothing marginals differ (later data informs them), filtering ones do not.
        assert np.array_equal(full["filt_mean"][:12], part["filt_mean"])


def test_mle_output_scale_is_constant_and_rescales_covariances(oracle):
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 3)
    save_at = np.linspace(0, 5, 6)
    cfg_m = oracle.Config(nu=3, solver="mle", rtol=1e-4, atol=1e-6)
    cfg_s = oracle.Config(nu=3, solver="solver", rtol=1e-4, atol=1e-6)
    rm = oracle.solve_save_at(cfg_m, THETA, tc, save_at)
    rs = oracle.solve_save_at(cfg_s, THETA, tc, save_at)
    os_ = rm["output_scale"][:, 0]
    assert np.allclose(os_, os_[0]) and os_[0] > 0
    assert np.allclose(rm["mean"], rs["mean"], rtol=1e-12)
    assert np.allclose(rm["chol"], os_[0] * rs["chol"], rtol=1e-12, atol=1e-300)


def test_status_reports_max_attempts_and_nonfinite(oracle):
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    cfg = oracle.Config(rtol=1e-300, atol=1e-300, max_attempts=5)
    r = oracle.solve_save_at(cfg, THETA, tc, np.array([0.0, 1.0]))
    assert r["status"] == 1
    bad = tc.copy()
    bad[0, 0] = np.nan
    r = oracle.solve_save_at(oracle.Config(), THETA, bad, np.array([0.0, 1.0]))
    assert r["status"] == 2


def _lml_bruteforce(oracle, cfg, r, data, stds, idx=0):
    """log p(y_0..y_K) from the Markov sequence, assembled as one dense Gaussian; the data observe
    Taylor coefficient ``idx``."""
    K = len(r["t"])
    n = cfg.n
    d = cfg.d
    total = 0.0
    cols = d if cfg.fact == "isotropic" else 1
    for col in range(cols):
        def meanvec(x):
            x = x.reshape(n, -1)
            return x[:, col : col + 1]

        # joint over x_0..x_{K-1}: x_{K-1} ~ N(term); x_j = M_j x_{j+1} + c_j + e_j
        dim = K * n
        mu = np.zeros(dim)
        Mfull = np.zeros((dim, dim))  # x = Mfull @ z + mu, z ~ N(0, I)
        L = r["term_chol"].reshape(n, n)
        mu[(K - 1) * n :] = meanvec(r["term_mean"])[:, 0]
        Mfull[(K - 1) * n :, (K - 1) * n :] = L
        for j in range(K - 2, -1, -1):
            A = r["cond_A"][j].reshape(n, n)
            nm = meanvec(r["cond_mean"][j])[:, 0]
            nc = r["cond_chol"][j].reshape(n, n)
            tl, to = r["cond_tl"][j], r["cond_to"][j]
            Mj = to[:, None] * A * tl[None, :]
            sl = slice(j * n, (j + 1) * n)
            sl1 = slice((j + 1) * n, (j + 2) * n)
            mu[sl] = Mj @ mu[sl1] + to * nm
            Mfull[sl, :] = Mj @ Mfull[sl1, :]
            Mfull[sl, sl] += np.abs(to)[:, None] * nc
        if cfg.fact == "dense":
            H = np.zeros((K * d, dim))
            for j in range(K):
                for k in range(d):
                    H[j * d + k, j * n + idx * d + k] = 1.0
            y = data.reshape(-1)
            R = np.diag(np.repeat(stds, d) ** 2) if np.ndim(stds) == 1 else np.diag(stds.reshape(-1) ** 2)
        else:
            H = np.zeros((K, dim))
            for j in range(K):
                H[j, j * n + idx] = 1.0
            y = data[:, col]
            R = np.diag(stds**2)
        cov = H @ Mfull @ Mfull.T @ H.T + R
        diff = y - H @ mu
        sign, logdet = np.linalg.slogdet(2 * np.pi * cov)
        total += -0.5 * (logdet + diff @ np.linalg.solve(cov, diff))
    return total


@pytest.mark.parametrize("solver", ["solver", "mle"])
@pytest.mark.parametrize("idx", [0, 1, 2])
@pytest.mark.parametrize("fact", ["isotropic", "dense"])
def test_lml_equals_bruteforce_joint_density(oracle, fact, idx, solver):
    """loss_lml_timeseries(tcoeff_index=idx) (estimators_and_losses.py:53-105) equals the joint
    Gaussian density of the data under the Markov sequence, with H selecting coefficient idx
    (solver_mle: the sequence its userfriendly_output rescales)."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 2)
    grid = np.linspace(0.0, 2.0, 41)
    every = 10
    cfg = oracle.Config(fact=fact, nu=2, strategy="fixedpoint", solver=solver, tcoeff_index=idx)
    r = oracle.solve_fixed_grid_checkpoints(cfg, THETA, tc, grid, every)
    K = len(r["t"])
    ref = lv_reference(grid[::every])
    rng = np.random.default_rng(4)
    data = ref + 0.1 * rng.normal(size=ref.shape)
    stds = np.full(K, 0.3)
    lml_sum = oracle.lml(cfg, THETA, tc, grid, every, data, stds, average=False)
    lml_avg = oracle.lml(cfg, THETA, tc, grid, every, data, stds, average=True)
    brute = _lml_bruteforce(oracle, cfg, r, data, stds, idx)
    assert np.isclose(lml_sum, brute, rtol=1e-8)
    assert np.isclose(lml_avg * K, lml_sum, rtol=1e-12)


@pytest.mark.parametrize("correct", [True, False])
@pytest.mark.parametrize("fact", ["isotropic", "dense"])
def test_lml_with_solver_mle_is_the_lml_at_the_calibrated_scale(oracle, fact, correct):
    """solver_mle (solvers.py:395-480) rescales the posterior by its calibrated output scale s.
    With noise-free ts0 observations and an exact initial state the means do not depend on the
    prior's scale and every covariance scales with s^2, so the LML of the rescaled sequence is the
    plain solver's LML under a prior of output scale s -- and differs from it at scale 1."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 3)
    grid = np.linspace(0.0, 3.0, 13)  # coarse, so that the solver's own uncertainty shows
    every = 2
    data = lv_reference(grid[::every]) + 0.05
    stds = np.full(data.shape[0], 0.01)
    mle = oracle.Config(fact=fact, nu=3, strategy="fixedpoint", solver="mle", correct_asymptotic_underconfidence=correct)
    r = oracle.solve_fixed_grid_checkpoints(mle, THETA, tc, grid, every)
    scale = float(np.ravel(r["output_scale"])[-1])
    assert scale > 0 and abs(scale - 1.0) > 1e-3
    plain = oracle.Config(fact=fact, nu=3, strategy="fixedpoint", solver="solver", base_scale=scale)
    got = oracle.lml(mle, THETA, tc, grid, every, data, stds, average=False)
    want = oracle.lml(plain, THETA, tc, grid, every, data, stds, average=False)
    assert np.isclose(got, want, rtol=1e-10)
    unit = oracle.lml(oracle.Config(fact=fact, nu=3, strategy="fixedpoint", solver="solver"), THETA, tc, grid, every,
                      data, stds, average=False)
    assert not np.isclose(got, unit, rtol=1e-3)


def test_calibration_tape_replays_and_freezes_the_output_scales(oracle):
    """The oracle's calibration tape (stop_gradient of the dynamic calibration, solvers.py:568-569):
    replaying the recorded scales reproduces the value bitwise; a wrong tape changes it; and the
    derivative with the scales frozen differs from the one through them (the calibration does
    depend on the parameters), by the small amount that sensitivity explains."""
    import lml_cases as LC

    params, y0, grid, every, data, stds = LC.lml_inputs(1, seed=3, t_end=5.0, n_grid=251, every=10)
    cfg = oracle.Config(fact="isotropic", nu=3, d=2, strategy="fixedpoint", solver="dynamic")
    p, y, dat, sd = params[0], y0[0], data[0], stds[0]
    tape = np.zeros(grid.size - 1)
    v_rec = oracle.lml_from_y0_tape(cfg, p, y, grid, every, dat, sd, tape, True)
    assert v_rec == oracle.lml_from_y0(cfg, p, y, grid, every, dat, sd)
    assert np.all(tape > 0)
    assert oracle.lml_from_y0_tape(cfg, p, y, grid, every, dat, sd, tape, False) == v_rec
    wrong = 2.0 * tape
    assert oracle.lml_from_y0_tape(cfg, p, y, grid, every, dat, sd, wrong, False) != v_rec
    with pytest.raises(ValueError, match="too short"):
        oracle.lml_from_y0_tape(cfg, p, y, grid, every, dat, sd, np.zeros(3), True)
    # d/d a with the scales free vs frozen (central differences, h = 1e-5)
    h = 1e-5
    e = np.zeros(4)
    e[0] = h
    free = (oracle.lml_from_y0(cfg, p + e, y, grid, every, dat, sd) - oracle.lml_from_y0(cfg, p - e, y, grid, every, dat, sd)) / (2 * h)
    frozen = (oracle.lml_from_y0_tape(cfg, p + e, y, grid, every, dat, sd, tape, False) - oracle.lml_from_y0_tape(cfg, p - e, y, grid, every, dat, sd, tape, False)) / (2 * h)
    assert free != frozen
    assert abs(free - frozen) < 1e-4 * abs(frozen)


@pytest.mark.parametrize("fact,strategy", [("dense", "filter"), ("isotropic", "fixedpoint")])
def test_re_linearize_after_calibration_changes_only_the_rounding(oracle, fact, strategy):
    """solvers.py:576-582: re-linearising at the calibrated prediction uses the same mean as the
    calibration's extrapolation (the mean does not depend on the output scale), so the solve is the
    same up to rounding: identical step counts, and outputs that differ from the plain solve by no
    more than rounding does, i.e. within 32x the spread of the two FP64 builds around the long
    double solve (the parity criterion of the device tests)."""
    import dataclasses

    cfg = oracle.Config(fact=fact, nu=4, d=2, strategy=strategy, solver="dynamic")
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 4)
    save_at = np.linspace(0.0, 30.0, 40)
    runs = {}
    for build in ("strict", "fma", "extended"):
        with oracle.rounding(build):
            runs[build] = oracle.solve_save_at(cfg, THETA, tc, save_at)
    r1 = oracle.solve_save_at(dataclasses.replace(cfg, re_linearize_after_calibration=True), THETA, tc, save_at)
    assert runs["strict"]["attempts"] == r1["attempts"]
    assert np.array_equal(runs["strict"]["num_steps"], r1["num_steps"])

    def rel(a, b):
        a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
        return np.max(np.abs(a - b)) / np.max(np.abs(b))

    for key in ("mean", "chol", "output_scale"):
        truth = runs["extended"][key]
        e_ref = max(rel(runs["strict"][key], truth), rel(runs["fma"][key], truth))
        assert rel(r1[key], truth) <= max(1e-10, 32 * e_ref), key


@pytest.mark.parametrize("fact", ["dense", "isotropic"])
def test_init_std_rows_place_the_initial_stds(oracle, fact):
    """from_mean_and_std with a std per covariance row (element-wise is_exact, diffuse_derivatives,
    ssm_impl_isotropic.py:409-447): the solution at t0 is the prior itself, so its covariance is
    diag(rows^2) in the state's row order (dense derivative-major); equal rows reproduce the scalar
    init_std bitwise."""
    tc = oracle.taylor_coefficients("lotka_volterra", THETA, [20.0, 20.0], 0.0, 3)
    n = 8 if fact == "dense" else 4
    rows = np.linspace(0.0, 0.7, n)
    times = np.linspace(0.0, 2.0, 5)
    cfg = oracle.Config(fact=fact, nu=3, d=2, strategy="filter", solver="solver", init_std_rows=rows)
    r = oracle.solve_save_at(cfg, THETA, tc, times)
    L0 = r["chol"][0].reshape(n, n)
    assert np.array_equal(L0 @ L0.T, np.diag(rows**2))
    same = oracle.solve_save_at(oracle.Config(fact=fact, nu=3, d=2, strategy="filter", solver="solver", init_std_rows=np.full(n, 0.3)), THETA, tc, times)
    ref = oracle.solve_save_at(oracle.Config(fact=fact, nu=3, d=2, strategy="filter", solver="solver"), THETA, tc, times, init_std=0.3)
    assert np.array_equal(same["mean"], ref["mean"]) and np.array_equal(same["chol"], ref["chol"])
