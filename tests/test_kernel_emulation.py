"""The kernels' per-instance code, built for the CPU, against the oracle.

``tests/emu/pdq_emu.cpp`` compiles the product's device sources with g++
(-DPDQ_HOST_EMU, one lane per group, one thread per workgroup).  This pins
the kernel logic -- loops, interpolation, smoother bookkeeping, status
codes -- on the CPU; the GPU tests pin the same code executed cooperatively
on gfx950.
"""

import numpy as np
import pytest

import test_gpu_parity as T  # noqa: E402  (shared comparison helpers)

pytestmark = pytest.mark.filterwarnings("ignore")


@pytest.fixture(scope="module")
def emu():
    from emu import emu as E

    E.build()
    E._lib = None
    return E


def _run(oracle, emu, cfg, params, y0s, times, fixed_grid=False, kernel="coop"):
    from probdiffeq_amd import capi

    B, K = times.shape
    ocfg = T._oracle_cfg(oracle, cfg)
    tcs = np.stack([oracle.taylor_coefficients(cfg.vector_field, params[i], y0s[i], times[i, 0], cfg.num_derivatives) for i in range(B)])
    g = emu.solve(cfg, params, tcs, times, fixed_grid, kernel=kernel)
    for i in range(B):
        r, rf, rt = T._oracle_solve(oracle, ocfg, params[i], tcs[i], times[i], fixed_grid)
        assert g["status"][i] == 0
        assert g["attempts"][i] == r["attempts"]
        assert np.array_equal(g["num_steps"][i], r["num_steps"])
        assert T._rel(g["t"][i], r["t"]) < 1e-14
        got = {k: g[k][i] for k in ("mean", "chol", "output_scale", "filt_mean", "filt_chol", "cond_A", "cond_mean", "term_mean") if k in g}
        T._check_fields(cfg, T._fields(cfg, got), r, rf, rt, f"emulation instance {i}")
    return g


def test_c1(oracle, emu):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="dense", strategy="filter", solver="dynamic")
    _run(oracle, emu, cfg, T.THETA[None], np.array([[20.0, 20.0]]), np.linspace(0, 50, 200)[None])


@pytest.mark.parametrize("fact", ["dense", "isotropic"])
@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
@pytest.mark.parametrize("solver", ["solver", "dynamic", "mle"])
def test_strategy_solver_grid(oracle, emu, fact, strategy, solver):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation=fact, num_derivatives=3, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7)
    p, y, t = T._lv_batch(4, seed=7, t_lo=3.0, t_hi=8.0, K=9)
    _run(oracle, emu, cfg, p, y, t)


def test_c3_ts1(oracle, emu):
    from probdiffeq_amd import capi

    rng = np.random.default_rng(3)
    mu = np.exp(rng.uniform(np.log(1.0), np.log(100.0), size=(3, 1)))
    cfg = capi.Config(factorisation="dense", num_derivatives=5, constraint="ts1", vector_field="van_der_pol", n_params=1, rtol=1e-5, atol=1e-7)
    _run(oracle, emu, cfg, mu, np.tile([2.0, 0.0], (3, 1)), np.tile([0.0, 6.3], (3, 1)))


@pytest.mark.parametrize("fact", ["dense", "isotropic"])
def test_fixed_grid(oracle, emu, fact):
    from probdiffeq_amd import capi

    p, y, _ = T._lv_batch(3, seed=11)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, strategy="fixedinterval", solver="dynamic")
    _run(oracle, emu, cfg, p, y, np.tile(np.linspace(0.0, 4.0, 41), (3, 1)), fixed_grid=True)


@pytest.mark.parametrize("solver", ["solver", "dynamic"])
def test_blockdiag_lotka_volterra(oracle, emu, solver):
    from probdiffeq_amd import capi

    p, y, t = T._lv_batch(3, seed=5)
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=4, strategy="filter", solver=solver)
    _run(oracle, emu, cfg, p, y, t)


def test_blockdiag_burgers(oracle, emu):
    from probdiffeq_amd import capi

    Ng = 64
    rng = np.random.default_rng(0)
    x = np.arange(Ng) / Ng
    y0 = np.stack([sum(rng.normal(0, 1 / k) * np.sin(2 * np.pi * k * x) for k in range(1, 9)) for _ in range(2)])
    params = np.tile([0.01, 1.0 / Ng], (2, 1))
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=2, ode_dim=Ng, vector_field="burgers", n_params=2, rtol=1e-4, atol=1e-6)
    _run(oracle, emu, cfg, params, y0, np.tile([0.0, 0.2], (2, 1)))


def _burgers(Ng, B, seed=0, nmodes=8):
    """The C4 initial condition at Ng points: sum of 8 sinusoids with N(0, 1/k^2) amplitudes."""
    rng = np.random.default_rng(seed)
    x = np.arange(Ng) / Ng
    y0 = np.stack([sum(rng.normal(0, 1 / k) * np.sin(2 * np.pi * k * x) for k in range(1, nmodes + 1)) for _ in range(B)])
    return np.tile([0.01, 1.0 / Ng], (B, 1)), y0


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
@pytest.mark.parametrize("solver", ["solver", "dynamic", "mle"])
def test_blockdiag_strategy_solver_grid(oracle, emu, strategy, solver):
    """The block-diagonal kernel (one lane per block, a group per instance) for every strategy and
    solver: Lotka-Volterra, d = 2 (G = 2), nu = 3."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="blockdiag", num_derivatives=3, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7)
    p, y, t = T._lv_batch(3, seed=17, t_lo=3.0, t_hi=8.0, K=9)
    _run(oracle, emu, cfg, p, y, t)


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint"])
def test_blockdiag_multi_wave_group(oracle, emu, strategy):
    """d = 100 > 64: one group of G = 128 lanes per instance (two wavefronts' worth, 28 padding
    lanes), workgroup barriers and the two-level sum; Burgers, nu = 2."""
    from probdiffeq_amd import capi

    params, y0 = _burgers(100, 2, seed=3)
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=2, ode_dim=100, vector_field="burgers", n_params=2, strategy=strategy, rtol=1e-4, atol=1e-6)
    _run(oracle, emu, cfg, params, y0, np.tile([0.0, 0.05, 0.1], (2, 1)))


@pytest.mark.parametrize("strategy", ["filter", "fixedinterval"])
def test_blockdiag_fixed_grid(oracle, emu, strategy):
    from probdiffeq_amd import capi

    p, y, _ = T._lv_batch(2, seed=23)
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=4, strategy=strategy, solver="dynamic")
    _run(oracle, emu, cfg, p, y, np.tile(np.linspace(0.0, 4.0, 41), (2, 1)), fixed_grid=True)


@pytest.mark.parametrize(
    "vf,d,strategy",
    [("logistic", 5, "filter"), ("logistic", 5, "fixedpoint"), ("lotka_volterra", 2, "filter"), ("burgers", 16, "filter")],
)
def test_blockdiag_ts1(oracle, emu, vf, d, strategy):
    """Blockdiag ts1 with the exact Jacobian diagonal (BlockDiagResidual.linearize,
    ssm_impl_blockdiag.py:147-183, jacobians.py:108-113)."""
    from probdiffeq_amd import capi

    rng = np.random.default_rng(29)
    B = 2
    if vf == "logistic":
        params = np.full((B, 1), 1.5)
        y0 = rng.uniform(0.1, 0.9, size=(B, d))
        times = np.tile(np.linspace(0.0, 4.0, 6), (B, 1))
        n_params = 1
    elif vf == "burgers":
        params, y0 = _burgers(d, B, seed=5, nmodes=3)
        times = np.tile([0.0, 0.1], (B, 1))
        n_params = 2
    else:
        params, y0, times = T._lv_batch(B, seed=31, t_lo=3.0, t_hi=6.0, K=7)
        n_params = 4
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=3, ode_dim=d, vector_field=vf, n_params=n_params, constraint="ts1", strategy=strategy, rtol=1e-5, atol=1e-7)
    _run(oracle, emu, cfg, params, y0, times)


@pytest.mark.parametrize(
    "solver,kw",
    [
        ("dynamic", dict(clip_dt=True)),
        ("dynamic", dict(error_norm="rms_then_scale")),
        ("dynamic", dict(control="proportional_integral")),
        ("dynamic", dict(error_estimator="state_std", error_derivative_idx=2)),
        ("solver", dict(error_estimator="state_std", error_re_linearize=True)),
        ("mle", dict(error_re_linearize=True, error_per_unit_step=True)),
        ("dynamic", dict(re_linearize_after_calibration=True)),
    ],
)
def test_blockdiag_options(oracle, emu, solver, kw):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="blockdiag", num_derivatives=3, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7, **kw)
    p, y, t = T._lv_batch(2, seed=37, t_lo=3.0, t_hi=6.0, K=7)
    _run(oracle, emu, cfg, p, y, t)


def test_status_codes(emu):
    from oracle import oracle as O
    from probdiffeq_amd import capi

    p, y, t = T._lv_batch(2, seed=1)
    tcs = np.stack([O.taylor_coefficients("lotka_volterra", p[i], y[i], 0.0, 4) for i in range(2)])
    g = emu.solve(capi.Config(rtol=1e-300, atol=1e-300, max_attempts=4), p, tcs, t)
    assert g["status"].tolist() == [1, 1]
    bad = tcs.copy()
    bad[0, 1, 0] = np.nan
    g = emu.solve(capi.Config(), p, bad, t)
    assert g["status"].tolist() == [2, 0]


@pytest.mark.parametrize(
    "fact,solver,per_coordinate_std,average",
    [("isotropic", "solver", False, True), ("isotropic", "dynamic", False, False), ("dense", "solver", True, True)],
)
def test_lml_value_and_gradient(oracle, emu, fact, solver, per_coordinate_std, average):
    """The LML kernel code: value against the oracle, forward-mode gradient against Richardson FD."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(2, seed=3, t_end=10.0, n_grid=501, every=10, per_coordinate_std=per_coordinate_std)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver=solver)
    ocfg = T._oracle_cfg(oracle, cfg)
    v, g = emu.lml(cfg, params, y0, grid, every, data, stds, average=average)
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lml instance {i}")
        gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        assert LC.rel(g[i], gr) < LC.GRAD_RTOL, (g[i], gr)


@pytest.mark.parametrize("nu,solver,average", [(3, "solver", True), (3, "dynamic", False), (4, "solver", True)])
def test_lml_reverse_adjoint(oracle, emu, nu, solver, average):
    """The reverse-mode adjoint of the lane LML (pdq_lane_adjoint.h: checkpoints every 50 grid steps,
    the forward factors recomputed per segment, the scan's and the steps' adjoints): its value is the
    forward LML's bitwise, its gradient equals the forward-mode tangents to rounding and the long
    double Richardson differences of the oracle to 1e-10.  Two scratch columns for three instances
    exercise the reuse of a lane's storage by the next instance."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=3, t_end=10.0, n_grid=501, every=10)
    cfg = capi.Config(factorisation="isotropic", num_derivatives=nu, ode_dim=2, strategy="fixedpoint", solver=solver)
    v_f, g_f = emu.lml(cfg, params, y0, grid, every, data, stds, average=average, kernel="lane")
    v_r, g_r = emu.lml_adjoint(cfg, params, y0, grid, every, data, stds, average=average, lanes=2)
    assert np.array_equal(v_r, v_f)
    assert LC.rel(g_r, g_f) < 1e-13
    ocfg = T._oracle_cfg(oracle, cfg)
    for i in range(2):
        gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        assert LC.rel(g_r[i], gr) < LC.GRAD_RTOL, (g_r[i], gr)


@pytest.mark.parametrize("solver", ["solver", "dynamic"])
def test_lml_with_an_exponential_prior(oracle, emu, solver):
    """loss_lml_timeseries on a solve with a Matern prior (prior_matern, ssm_impl_api.py:443-477;
    the dense exponential transition every step) in the cooperative kernel's code: value against
    the oracle, forward-mode gradient against Richardson FD (the transition depends on the step
    and the drift, not on the parameters)."""
    import bench
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(2, seed=6, t_end=5.0, n_grid=101, every=5, per_coordinate_std=True)
    drift = bench.matern_drift(0.3, 4, 2)
    cfg = capi.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver=solver,
                      prior="exponential", prior_drift=drift)
    ocfg = T._oracle_cfg(oracle, cfg)
    v, g = emu.lml(cfg, params, y0, grid, every, data, stds)
    v_iwp, _ = emu.lml(capi.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver=solver), params, y0, grid, every, data, stds, grad=False)
    assert np.all(np.abs(v - v_iwp) > 1e-3 * np.abs(v_iwp))  # another prior: another likelihood
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lml matern {solver} instance {i}")
        gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
        assert LC.rel(g[i], gr) < LC.GRAD_RTOL, (g[i], gr)


def test_lml_gain_solves_with_a_dropped_singular_value(oracle, emu):
    """loss_lml_timeseries(solve_triu=linalg.lstsq_svd / linalg.solve_triu) in the cooperative
    kernel's code (dense, two observed rows per node).  Where R_Y's small singular value falls below
    lstsq_svd's cutoff (lml_dropped_singular_value_inputs) the kernel's Jacobi-SVD gain matches the
    oracle's lstsq_svd and its back substitution the oracle's solve_triu, value and gradient; the
    two solves' likelihoods differ by far more than the tolerance, so each comparison tells them
    apart."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_dropped_singular_value_inputs()
    vals = {}
    for solve in ("lstsq_svd", "solve_triu"):
        cfg = capi.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic", lml_solve=solve)
        ocfg = T._oracle_cfg(oracle, cfg)
        assert ocfg.lml_solve == solve
        v, g = emu.lml(cfg, params, y0, grid, every, data, stds)
        for i in range(2):
            vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
            T.check_accuracy(v[i], vo, vf, vt, f"emulation lml {solve} instance {i}")
            gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
            assert LC.rel(g[i], gr) < LC.GRAD_RTOL, (solve, g[i], gr)
        vals[solve] = v
    assert np.all(np.abs(vals["lstsq_svd"] - vals["solve_triu"]) > 1e-3 * np.abs(vals["solve_triu"]))


@pytest.mark.parametrize("fact,idx,grad", [("isotropic", 1, True), ("isotropic", 2, True), ("isotropic", 3, False), ("dense", 1, True), ("dense", 2, True), ("dense", 3, False)])
def test_lml_observing_a_derivative(oracle, emu, fact, idx, grad):
    """loss_lml_timeseries(tcoeff_index=idx) in the cooperative kernel's code: the data observe Taylor
    coefficient idx (to_derivative, ssm_impl_dense.py:206-217, ssm_impl_isotropic.py:271-278); value
    against the oracle, forward-mode gradient against Richardson FD.  For idx = nu the long double
    oracle's value moves by ~1e-13 relative under perturbations (the highest coefficient is the least
    well conditioned), so finite differences cannot pin the gradient to 1e-10: value only."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(2, seed=9, t_end=10.0, n_grid=501, every=10)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic", lml_tcoeff_index=idx)
    ocfg = T._oracle_cfg(oracle, cfg)
    assert ocfg.tcoeff_index == idx
    v, g = emu.lml(cfg, params, y0, grid, every, data, stds)
    v0, _ = emu.lml(capi.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic"), params, y0, grid, every, data, stds, grad=False)
    assert np.all(np.abs(v - v0) > 1.0)  # another coefficient observed: another likelihood
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lml tcoeff {idx} instance {i}")
        if grad:
            gr, tol = LC.grad_reference(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
            assert tol < 1e-8
            assert LC.rel(g[i], gr) < tol, (g[i], gr, tol)


@pytest.mark.parametrize("fact", ["isotropic", "dense"])
def test_lml_value_with_ts1(oracle, emu, fact):
    """The cooperative LML's value with ts1 (isotropic: the exact trace) against the oracle."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(2, seed=14, t_end=10.0, n_grid=501, every=10)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic", constraint="ts1")
    ocfg = T._oracle_cfg(oracle, cfg)
    v, _ = emu.lml(cfg, params, y0, grid, every, data, stds, grad=False)
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lml ts1 instance {i}")


@pytest.mark.parametrize("fact,correct,constraint", [("isotropic", True, "ts0"), ("dense", False, "ts0"), ("isotropic", True, "ts1")])
def test_lml_value_with_solver_mle(oracle, emu, fact, correct, constraint):
    """The cooperative LML's value with solver_mle: the posterior rescaled by the calibrated output
    scale (solvers.py:439-480) before the reverse scan, against the oracle."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs_scale_sensitive(2, seed=16)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="mle",
                      constraint=constraint, correct_asymptotic_underconfidence=correct)
    ocfg = T._oracle_cfg(oracle, cfg)
    v, _ = emu.lml(cfg, params, y0, grid, every, data, stds, grad=False)
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i])
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lml mle instance {i}")
        # the rescaling is live: the uncalibrated solver's LML is far off
        plain = oracle.lml_from_y0(T._oracle_cfg(oracle, capi.Config(**{**cfg.__dict__, "solver": "solver"})),
                                   params[i], y0[i], grid, every, data[i], stds[i])
        assert abs(plain - vo) > 1e-3 * abs(vo)


@pytest.mark.parametrize("fact,kernel", [("dense", "coop"), ("isotropic", "coop"), ("isotropic", "lane")])
def test_init_std_rows(oracle, emu, fact, kernel):
    """A std per covariance row at t0 (element-wise is_exact, diffuse_derivatives) in both kernels'
    code, against the oracle."""
    from probdiffeq_amd import capi

    n = 10 if fact == "dense" else 5
    rows = np.array([0.0, 0.0, 1e-3, 1e-3, 0.5, 0.5, 1.0, 1.0, 1.0, 1.0])[:n] if fact == "dense" else np.array([0.0, 0.0, 1e-3, 0.5, 1.0])
    cfg = capi.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver="dynamic", rtol=1e-5, atol=1e-7, init_std_rows=rows)
    ocfg = T._oracle_cfg(oracle, cfg)
    assert ocfg.init_std_rows is not None
    p, y, t = T._lv_batch(3, seed=21, t_lo=4.0, t_hi=8.0, K=10)
    _run(oracle, emu, cfg, p, y, t, kernel=kernel)


# ----------------------------------------------- lane-per-instance kernel
@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
@pytest.mark.parametrize("solver", ["solver", "dynamic", "mle"])
def test_lane_kernel_strategy_solver_grid(oracle, emu, strategy, solver):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=3, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7)
    p, y, t = T._lv_batch(4, seed=7, t_lo=3.0, t_hi=8.0, K=9)
    _run(oracle, emu, cfg, p, y, t, kernel="lane")


@pytest.mark.parametrize(
    "kw",
    [
        dict(strategy="fixedpoint", solver="dynamic"),
        dict(strategy="filter", solver="dynamic", clip_dt=True),
        dict(strategy="filter", solver="dynamic", error_norm="rms_then_scale"),
        dict(strategy="fixedpoint", solver="dynamic", control="proportional_integral"),
    ],
)
def test_lane_kernel_c2_shape(oracle, emu, kw):
    """The C2 shape (isotropic nu=4, d=2) and the configuration switches."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, **kw)
    p, y, t = T._lv_batch(3, seed=1, t_lo=5.0, t_hi=15.0, K=20)
    _run(oracle, emu, cfg, p, y, t, kernel="lane")


@pytest.mark.parametrize("strategy", ["filter", "fixedinterval"])
def test_lane_kernel_fixed_grid(oracle, emu, strategy):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=3, strategy=strategy, solver="dynamic")
    p, y, _ = T._lv_batch(3, seed=11)
    times = np.tile(np.linspace(0.0, 4.0, 41), (3, 1))
    _run(oracle, emu, cfg, p, y, times, fixed_grid=True, kernel="lane")


def test_lane_kernel_status_codes(emu):
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, rtol=1e-300, atol=1e-300, max_attempts=4)
    p, y, t = T._lv_batch(2, seed=1)
    from oracle import oracle as O

    tcs = np.stack([O.taylor_coefficients("lotka_volterra", p[i], y[i], 0.0, 4) for i in range(2)])
    g = emu.solve(cfg, p, tcs, t, kernel="lane")
    assert g["status"].tolist() == [1, 1]


def test_lane_powers_of_dt_are_within_half_an_ulp(emu):
    """The lane kernel's dt^k (preconditioner_taylor, utilities.py:74-84, which calls pow):
    against a long-double pow, every result is within 0.5 ulp plus a small margin, as a
    correctly rounded pow is; special values follow pow."""
    import ctypes

    f = emu.lib().pdqemu_ipow
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int, ctypes.c_double]
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(np.log(1e-12), np.log(1e3), 4000)), [1e-3, 0.1, 0.3, 1.0, 2.0, 7.5]])
    xs = np.concatenate([xs, -xs[:500]])
    worst = 0.0
    for k in range(-7, 8):
        for x in xs:
            got = f(k, float(x))
            exact = np.longdouble(x) ** k
            if not np.isfinite(float(exact)) or float(exact) == 0.0:
                continue
            ulp = np.spacing(abs(float(exact)))
            err = abs(np.longdouble(got) - exact) / np.longdouble(ulp)
            worst = max(worst, float(err))
    assert worst < 0.52, worst
    assert f(-4, 0.0) == np.inf and f(-3, -0.0) == -np.inf
    assert f(4, 0.0) == 0.0


def test_lane_compile_time_transition_matches_the_host_matrices(emu):
    """The lane kernel's compile-time IWP transition equals the host's iwp_matrices (utilities.py:57-71)
    entry for entry, and the host's Q factor is lower triangular, for nu = 1..7."""
    assert emu.lib().pdqemu_iwp_mismatches() == 0


@pytest.mark.parametrize("rescale", [0, 1])
def test_lane_revert_halves_are_bitwise_the_revert(emu, rescale):
    """The fixed-point smoother's adaptive loop splits the revert (ssm_impl_isotropic.py:107-133)
    into the prediction (lane_revert_left) and the backward conditional (lane_revert_right, on the
    merging wavefront): together they give lane_revert_prior's prediction and conditional bit for
    bit, for nu = 1..5, two and four mean columns, with and without the identity merge's
    rescaling, including a zero factor (no reflections) and a factor with a zero column."""
    for n in range(2, 7):
        for d in (2, 4):
            assert emu.lib().pdqemu_revert_halves_mismatch(n, d, 200, 11 + n, rescale) == 0, (n, d)


def test_lane_shared_divisor_quotient_is_the_correctly_rounded_division(emu):
    """lane_div (one reciprocal per divisor, then an FMA correction) returns exactly a / b:
    random operands over a wide exponent range, hard cases near ties, and IEEE special values."""
    import ctypes

    f = emu.lib().pdqemu_lane_div
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(3)
    n = 200_000
    a = rng.uniform(1, 2, n) * 2.0 ** rng.integers(-300, 300, n) * rng.choice([-1, 1], n)
    b = rng.uniform(1, 2, n) * 2.0 ** rng.integers(-300, 300, n) * rng.choice([-1, 1], n)
    # operands whose quotient lies close to a rounding boundary: a = b * q(1 + tiny)
    q = rng.uniform(1, 2, 20_000)
    bb = rng.uniform(1, 2, 20_000)
    a = np.concatenate([a, bb * q, np.nextafter(bb * q, np.inf), np.nextafter(bb * q, -np.inf)])
    b = np.concatenate([b, bb, bb, bb])
    got = np.array([f(x, y) for x, y in zip(a.tolist(), b.tolist())])
    assert np.array_equal(got, a / b)
    specials = [(1.0, 0.0), (-1.0, 0.0), (0.0, 0.0), (np.inf, 2.0), (2.0, np.inf), (np.nan, 1.0), (1.0, 1e-310), (1e-310, 3.0), (0.0, 5.0), (-0.0, 5.0)]
    for x, y in specials:
        with np.errstate(all="ignore"):
            e = np.float64(x) / np.float64(y)
        g = f(x, y)
        assert (np.isnan(e) and np.isnan(g)) or (g == e and np.signbit(g) == np.signbit(e)), (x, y, g, e)


# ------------------------------------------------ lane-per-instance LML
@pytest.mark.parametrize(
    "nu,solver,average,tpl",
    [(3, "solver", True, 3), (3, "solver", True, 1), (3, "dynamic", False, 3), (4, "solver", True, 2), (4, "dynamic", True, 1)],
)
def test_lane_lml_value_and_gradient(oracle, emu, nu, solver, average, tpl):
    """The lane LML code (each lane: the shared factors, the primal and `tpl` tangent columns):
    value against the oracle, forward-mode gradient against Richardson FD of the oracle."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(2, seed=3, t_end=10.0, n_grid=501, every=10)
    cfg = capi.Config(factorisation="isotropic", num_derivatives=nu, ode_dim=2, strategy="fixedpoint", solver=solver)
    ocfg = T._oracle_cfg(oracle, cfg)
    v, g = emu.lml(cfg, params, y0, grid, every, data, stds, average=average, kernel="lane", tangents_per_lane=tpl)
    for i in range(2):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        T.check_accuracy(v[i], vo, vf, vt, f"emulation lane lml instance {i}")
        gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        assert LC.rel(g[i], gr) < LC.GRAD_RTOL, (g[i], gr)


def test_lane_lml_value_only_matches_the_gradient_run(oracle, emu):
    """Without tangents (tpl = 0) the lane LML gives the same value as with them."""
    import lml_cases as LC
    from probdiffeq_amd import capi

    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=8, t_end=10.0, n_grid=501, every=10)
    cfg = capi.Config(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver")
    v0, _ = emu.lml(cfg, params, y0, grid, every, data, stds, grad=False, kernel="lane")
    v1, _ = emu.lml(cfg, params, y0, grid, every, data, stds, kernel="lane")
    assert np.array_equal(v0, v1)


def _matern_drift(length_scale, N, d):
    """prior_matern's drift (ssm_impl_api.py:443-477), pinned in tests/test_oracle_prior.py."""
    from math import comb

    z = np.sqrt(2 * (N - 0.5)) / length_scale
    return np.concatenate([-comb(N, i) * z ** (N - i) * np.eye(d) for i in range(N)], axis=1)


# ------------------------------------------------ off-grid marginals
@pytest.mark.parametrize("nu,strategy,solver", [(4, "filter", "dynamic"), (3, "filter", "solver"), (4, "fixedinterval", "dynamic"), (3, "fixedinterval", "mle")])
def test_offgrid_marginals_kernel_code(oracle, emu, nu, strategy, solver):
    """The off-grid marginal kernel's code (pdq_lane_offgrid.h) on an oracle solution, against the
    oracle's offgrid_marginal (itself pinned to the Gaussian closed forms) in its three builds."""
    import offgrid_cases as OC

    cfg, times, sol, tq, refs = OC.make_case(oracle, nu, strategy, solver)
    batch = {k: v[None] for k, v in sol.items()}
    m, L = emu.offgrid(cfg, times, batch, tq)
    OC.check(m[0], L[0], refs, f"emulation offgrid {strategy}/{solver} nu={nu}")


@pytest.mark.parametrize(
    "fact,nu,strategy,solver,kw",
    [
        ("dense", 4, "filter", "dynamic", {}),
        ("dense", 3, "fixedinterval", "dynamic", {}),
        ("dense", 3, "fixedinterval", "solver", {}),
        ("isotropic", 2, "fixedinterval", "dynamic", {}),
        ("isotropic", 4, "filter", "dynamic", dict(lanes_per_instance=16)),
        ("dense", 3, "fixedinterval", "dynamic", dict(prior="exponential", prior_drift=_matern_drift(8.0, 4, 2))),
    ],
)
def test_offgrid_marginals_cooperative_code(oracle, emu, fact, nu, strategy, solver, kw):
    """Off-grid marginals on the cooperative kernel's primitives (dense; isotropic shapes without a
    lane kernel; exponential priors), against the oracle's offgrid_marginal in its three builds."""
    import offgrid_cases as OC

    cfg, times, sol, tq, refs = OC.make_case(oracle, nu, strategy, solver, fact=fact, **kw)
    batch = {k: v[None] for k, v in sol.items()}
    m, L = emu.offgrid(cfg, times, batch, tq)
    OC.check(m[0], L[0], refs, f"emulation coop offgrid {fact}/{strategy}/{solver} nu={nu}", cov=fact == "dense")


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint"])
def test_lane_kernel_re_linearize_after_calibration(oracle, emu, strategy):
    """solver_dynamic(re_linearize_after_calibration=True) (solvers.py:576-582) in the lane kernel's code."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy=strategy, solver="dynamic", re_linearize_after_calibration=True)
    p, y, t = T._lv_batch(3, seed=2, t_lo=5.0, t_hi=10.0, K=15)
    _run(oracle, emu, cfg, p, y, t, kernel="lane")


@pytest.mark.parametrize("nu", [3, 4])
def test_posterior_sampler_kernel_code(oracle, emu, nu):
    """The sampler's code (pdq_lane_sample.h) against the oracle's MarkovSequence.sample with the
    same draws (the oracle pinned in tests/test_oracle_sample.py)."""
    import sample_cases as SC

    cfg, sol, base, refs = SC.make_case(oracle, nu)
    SC.check(emu.sample(cfg, sol, base), refs, f"emulation sample nu={nu}")


@pytest.mark.parametrize(
    "solver,kw",
    [
        ("dynamic", dict(error_estimator="state_std")),
        ("dynamic", dict(error_estimator="state_std", error_derivative_idx=2, error_per_unit_step=True)),
        ("solver", dict(error_estimator="state_std", error_re_linearize=True)),
        ("mle", dict(error_estimator="residual_std", error_re_linearize=True, error_per_unit_step=True)),
        ("dynamic", dict(error_estimator="residual_std", error_per_unit_step=True)),
    ],
)
def test_lane_kernel_error_estimators(oracle, emu, solver, kw):
    """error_state_std (solvers.py:999-1098) and the estimators' options in the lane kernel's code."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7, **kw)
    p, y, t = T._lv_batch(3, seed=5, t_lo=4.0, t_hi=8.0, K=10)
    _run(oracle, emu, cfg, p, y, t, kernel="lane")


@pytest.mark.parametrize(
    "strategy,solver,vf,kw",
    [
        ("fixedpoint", "dynamic", "lotka_volterra", {}),
        ("filter", "solver", "lotka_volterra", {}),
        ("fixedinterval", "mle", "lotka_volterra", {}),
        ("fixedpoint", "dynamic", "lotka_volterra", dict(error_estimator="state_std", error_derivative_idx=1)),
        ("fixedpoint", "dynamic", "lotka_volterra", dict(re_linearize_after_calibration=True)),
        ("filter", "solver", "lotka_volterra", dict(error_re_linearize=True)),
        ("filter", "dynamic", "van_der_pol", {}),
    ],
)
@pytest.mark.parametrize("kernel", ["lane", "coop"])
def test_isotropic_ts1_kernels(oracle, emu, strategy, solver, vf, kw, kernel):
    """Isotropic ts1 with the exact trace (ssm_impl_isotropic.py:321-355, jacobians.py:100-106) in
    the lane kernel's and the cooperative kernel's code (the oracle pinned by the dense == isotropic
    equivalence)."""
    from probdiffeq_amd import capi

    if vf == "van_der_pol":
        rng = np.random.default_rng(4)
        p = np.exp(rng.uniform(np.log(1.0), np.log(10.0), size=(3, 1)))
        y, t = np.tile([2.0, 0.0], (3, 1)), np.tile(np.linspace(0.0, 6.3, 8), (3, 1))
        cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy=strategy, solver=solver, constraint="ts1",
                          vector_field=vf, n_params=1, rtol=1e-5, atol=1e-7, **kw)
    else:
        p, y, t = T._lv_batch(3, seed=6, t_lo=4.0, t_hi=8.0, K=10)
        cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy=strategy, solver=solver, constraint="ts1",
                          rtol=1e-5, atol=1e-7, **kw)
    _run(oracle, emu, cfg, p, y, t, kernel=kernel)


@pytest.mark.parametrize(
    "fact,kernel,strategy,kw",
    [
        ("isotropic", "lane", "fixedpoint", {}),
        ("isotropic", "coop", "fixedpoint", {}),
        ("isotropic", "lane", "filter", dict(re_linearize_after_calibration=True)),
        ("isotropic", "coop", "fixedinterval", dict(re_linearize_after_calibration=True)),
        ("blockdiag", "coop", "fixedpoint", {}),
        ("blockdiag", "coop", "filter", dict(re_linearize_after_calibration=True, clip_dt=True)),
    ],
)
@pytest.mark.parametrize("mode", ["monte_carlo_rev", "monte_carlo_fwd"])
def test_monte_carlo_jacobian_kernels(oracle, emu, fact, kernel, strategy, kw, mode):
    """ts1 with the Monte Carlo Jacobian handlers (jacobians.py:116-271): the kernels draw the
    oracle's probes (one set per linearisation, indexed by the accepted steps) and follow its
    solves, rejected attempts included."""
    from probdiffeq_amd import capi

    p, y, t = T._lv_batch(3, seed=8, t_lo=4.0, t_hi=8.0, K=10)
    cfg = capi.Config(factorisation=fact, num_derivatives=3, strategy=strategy, solver="dynamic", constraint="ts1",
                      rtol=1e-5, atol=1e-7, jacobian=mode, jac_probes=6, jac_seed=3, **kw)
    g = _run(oracle, emu, cfg, p, y, t, kernel=kernel)
    assert np.all(g["attempts"] > g["accepted"])  # rejected attempts reuse their probe sets
    fx = _run(oracle, emu, cfg, p, y, t, fixed_grid=True, kernel=kernel)
    assert np.all(fx["status"] == 0)


def test_monte_carlo_jacobian_blockdiag_burgers(oracle, emu):
    """Blockdiag ts1 on Burgers (three non-zeros per Jacobian column, wrapped at the ends) with
    the reverse-mode diagonal estimate."""
    from probdiffeq_amd import capi

    params, y0 = _burgers(16, 2, seed=5, nmodes=3)
    times = np.tile([0.0, 0.05, 0.1], (2, 1))
    cfg = capi.Config(factorisation="blockdiag", num_derivatives=3, ode_dim=16, vector_field="burgers", n_params=2,
                      constraint="ts1", strategy="filter", rtol=1e-5, atol=1e-7, jacobian="monte_carlo_rev", jac_probes=10)
    _run(oracle, emu, cfg, params, y0, times)


@pytest.mark.parametrize(
    "prior,strategy,solver",
    [
        ("zero", "filter", "dynamic"),
        ("matern", "filter", "dynamic"),
        ("ioup", "fixedpoint", "dynamic"),
        ("matern", "fixedinterval", "solver"),
        ("ioup", "filter", "mle"),
    ],
)
def test_coop_kernel_exponential_prior(oracle, emu, prior, strategy, solver):
    """The exponential prior (ssm_impl_dense.py:392-629, exp_gram_cholesky) in the cooperative
    kernel's code, against the oracle pinned in tests/test_oracle_prior.py."""
    from probdiffeq_amd import capi

    nu, d = 3, 2
    N, n = nu + 1, (nu + 1) * 2
    if prior == "zero":
        drift = np.zeros((d, n))
    elif prior == "matern":
        drift = _matern_drift(8.0, N, d)
    else:
        drift = np.zeros((d, n))
        drift[:, nu * d :] = np.array([[-0.3, 0.1], [0.05, -0.2]])
    cfg = capi.Config(factorisation="dense", num_derivatives=nu, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7,
                      prior="exponential", prior_drift=drift)
    p, y, t = T._lv_batch(3, seed=8, t_lo=4.0, t_hi=8.0, K=10)
    _run(oracle, emu, cfg, p, y, t)


@pytest.mark.parametrize("fact,kernel", [("isotropic", "lane"), ("dense", "coop")])
def test_per_instance_dt0(oracle, emu, fact, kernel):
    """dt0_per_instance: each instance starts from its own dt0 (a vmap of the reference's solve over
    per-instance dt0), exactly as the oracle run with that instance's dt0."""
    from probdiffeq_amd import capi

    p, y, t = T._lv_batch(3, seed=9, t_lo=3.0, t_hi=6.0, K=6)
    dt0s = np.array([0.01, 0.3, 0.07])
    cfg = capi.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver="dynamic", rtol=1e-5, atol=1e-7, dt0=0.5,
                      dt0_per_instance=dt0s)
    tcs = np.stack([oracle.taylor_coefficients("lotka_volterra", p[i], y[i], 0.0, 4) for i in range(3)])
    g = emu.solve(cfg, p, tcs, t, kernel=kernel)
    for i in range(3):
        cfg_i = capi.Config(**{**cfg.__dict__, "dt0": float(dt0s[i]), "dt0_per_instance": None})
        r, rf, rt = T._oracle_solve(oracle, T._oracle_cfg(oracle, cfg_i), p[i], tcs[i], t[i], False)
        assert g["attempts"][i] == r["attempts"]
        assert np.array_equal(g["num_steps"][i], r["num_steps"])
        got = {k: g[k][i] for k in ("mean", "chol", "output_scale", "filt_mean", "filt_chol", "cond_A", "cond_mean", "term_mean") if k in g}
        T._check_fields(cfg_i, T._fields(cfg_i, got), r, rf, rt, f"per-instance dt0 {kernel} instance {i}")
    r0 = oracle.solve_save_at(T._oracle_cfg(oracle, capi.Config(**{**cfg.__dict__, "dt0_per_instance": None})), p[1], tcs[1], t[1])
    assert r0["attempts"] != g["attempts"][1] or not np.allclose(r0["mean"], g["mean"][1], rtol=1e-12)


@pytest.mark.parametrize(
    "fact,nu,strategy,kw",
    [
        ("dense", 4, "fixedpoint", {}),
        ("dense", 3, "fixedinterval", {}),
        ("isotropic", 2, "fixedpoint", {}),
        ("isotropic", 4, "fixedinterval", dict(lanes_per_instance=16)),
        ("dense", 3, "fixedpoint", dict(prior="exponential", prior_drift=_matern_drift(8.0, 4, 2))),
    ],
)
def test_generic_sampler_code(oracle, emu, fact, nu, strategy, kw):
    """The sampler for the shapes without a lane sampler (pdq_sample_generic.h): dense (to applied to
    the whole mean, n draws per step) and isotropic (the reference's isotropic forms), against the
    oracle's MarkovSequence.sample (pinned in tests/test_oracle_sample.py) in its three builds."""
    import sample_cases as SC

    cfg, sol, base, refs = SC.make_case(oracle, nu, fact=fact, strategy=strategy, **kw)
    SC.check(emu.sample(cfg, sol, base), refs, f"emulation generic sample {fact}/{strategy} nu={nu}")


@pytest.mark.parametrize(
    "fact,solver,kw",
    [
        ("dense", "dynamic", dict(error_derivative_idx=0)),
        ("dense", "dynamic", dict(error_derivative_idx=2, re_linearize_after_calibration=True)),
        ("dense", "solver", dict(error_derivative_idx=0, error_re_linearize=True)),
        ("dense", "mle", dict(error_derivative_idx=3, error_per_unit_step=True, error_norm="rms_then_scale")),
        ("isotropic", "dynamic", dict(error_derivative_idx=2)),
    ],
)
def test_cooperative_kernel_error_state_std(oracle, emu, fact, solver, kw):
    """error_state_std (solvers.py:999-1098) in the cooperative kernel's code: dense (d calibrated
    stds of derivative idx) and isotropic, against the oracle.  Derivative 1 is left out: the ODE
    constraint pins it, so its posterior std is rounding noise."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7,
                      error_estimator="state_std", **kw)
    p, y, t = T._lv_batch(3, seed=13, t_lo=4.0, t_hi=8.0, K=10)
    _run(oracle, emu, cfg, p, y, t)
