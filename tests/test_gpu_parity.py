"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Step counts (attempts, accepted steps, num_steps) and the output times are
identical.  Every floating-point output array X is compared normwise
(rel(a, b) = max |a - b| / max |b|) against three oracle solves of the same
float64 inputs: the FP64 restatement compiled with strict evaluation order
("ref") and with FMA contraction ("fma"), two valid FP64 roundings of the
reference's algorithm, and the same restatement in long double ("truth",
64-bit mantissa).  The FP64 reference's own rounding error on X is

    e_ref = max(rel(ref, truth), rel(fma, truth)),

and the kernels must meet

    rel(gpu, truth) <= tol  and  rel(gpu, ref) <= tol,
    tol = max(1e-10, ACCURACY_FACTOR * e_ref).

Where X is well conditioned (the ODE solution u(t): e_ref ~ 1e-15) this is
1e-10 relative.  Some outputs are not: the top Taylor coefficients (scaled by
dt^-k), covariances and the dynamic calibration carry FP64 rounding errors
of 1e-12 .. 1e-3 in the reference itself (the IWP noise factor is the
Cholesky factor of a Hilbert-type matrix; on stiff van der Pol with ts1 the
calibration's whitened residual is at cancellation level); there the GPU
result must be as accurate as the FP64 reference within ACCURACY_FACTOR.
e_ref is estimated from two FP64 samples and the GPU's error is a third, so
their ratio scatters: over the 2454 comparisons of one run of this suite
(both kernels) its median is 1.1, its 99th percentile 4.7, its 99.9th
percentile 10.9 and its maximum 23.6, hence ACCURACY_FACTOR = 32.
Covariances are compared as L L^T, and the dense filter solve also compares
the factors themselves, up to column signs.  With
PDQ_PARITY_REPORT=<path>, every comparison is written there as JSON.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

THETA = np.array([0.5, 0.05, 0.5, 0.05])
RTOL_FLOOR = 1e-10
ACCURACY_FACTOR = 32.0
REPORT = []


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / scale


def _cov(chol, n, nblk=1):
    L = np.asarray(chol).reshape(-1, nblk, n, n)
    return np.einsum("kbij,kblj->kbil", L, L)


@pytest.fixture(scope="module")
def pdq():
    from probdiffeq_amd import capi

    capi.require_device()
    capi.lib()
    return capi


def _u0(cfg, mean):
    """Taylor coefficient 0 (the ODE solution) of flat means (..., mean_size)."""
    m = np.asarray(mean)
    N, d = cfg.N, cfg.ode_dim
    if cfg.factorisation == "blockdiag":
        return m.reshape(*m.shape[:-1], d, N)[..., 0]
    return m[..., :d]


def _cov0(cfg, cov):
    """Covariance block of Taylor coefficient 0 from (K, nblk, n, n) covariances."""
    if cfg.factorisation == "dense":
        return cov[..., : cfg.ode_dim, : cfg.ode_dim]
    return cov[..., :1, :1]


def _dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


def _oracle_cfg(oracle, cfg):
    return oracle.Config(
        fact=cfg.factorisation,
        nu=cfg.num_derivatives,
        d=cfg.ode_dim,
        constraint=cfg.constraint,
        strategy=cfg.strategy,
        solver=cfg.solver,
        error_norm=cfg.error_norm,
        control=cfg.control,
        vf=cfg.vector_field,
        clip_dt=cfg.clip_dt,
        rtol=cfg.rtol,
        atol=cfg.atol,
        dt0=cfg.dt0,
        eps=cfg.eps,
        damp=cfg.damp,
        base_scale=cfg.base_output_scale,
        max_attempts=cfg.max_attempts,
        correct_asymptotic_underconfidence=cfg.correct_asymptotic_underconfidence,
        re_linearize_after_calibration=cfg.re_linearize_after_calibration,
        error_estimator=cfg.error_estimator,
        error_re_linearize=cfg.error_re_linearize,
        error_derivative_idx=cfg.error_derivative_idx,
        error_per_unit_step=cfg.error_per_unit_step,
        prior=cfg.prior,
        prior_drift=None if cfg.prior_drift is None else np.asarray(cfg.prior_drift.cpu() if hasattr(cfg.prior_drift, "cpu") else cfg.prior_drift),
        tcoeff_index=cfg.lml_tcoeff_index,
        init_std_rows=None if cfg.init_std_rows is None else np.asarray(cfg.init_std_rows.cpu() if hasattr(cfg.init_std_rows, "cpu") else cfg.init_std_rows),
        lml_solve=cfg.lml_solve,
        jacobian=cfg.jacobian,
        jac_probes=cfg.jac_probes,
        jac_seed=cfg.jac_seed,
    )


def _fields(cfg, r):
    """The compared output arrays of one solve (covariances as L L^T)."""
    n, nblk = cfg.n, cfg.nblk
    cov = _cov(r["chol"], n, nblk)
    out = {"u0": _u0(cfg, r["mean"]), "mean": r["mean"], "cov0": _cov0(cfg, cov), "cov": cov, "output_scale": r["output_scale"]}
    if cfg.strategy != "filter":
        out.update(
            filt_mean=r["filt_mean"],
            filt_cov=_cov(r["filt_chol"], n, nblk),
            cond_A=r["cond_A"],
            cond_mean=r["cond_mean"],
            term_mean=r["term_mean"],
        )
    return out


def _oracle_solve(oracle, ocfg, params, tcs, times, fixed_grid):
    """(FP64 strict, FP64 FMA, long double) oracle solves of the same float64 inputs."""
    f = oracle.solve_fixed_grid if fixed_grid else oracle.solve_save_at
    ref = f(ocfg, params, tcs, times)
    with oracle.rounding("fma"):
        fma = f(ocfg, params, tcs, times)
    with oracle.rounding("extended"):
        truth = f(ocfg, params, tcs, times)
    for r in (ref, fma, truth):
        assert r["status"] == 0
        assert r["attempts"] == ref["attempts"], "the oracle builds took different step sequences"
    return ref, fma, truth


def check_accuracy(got, ref, fma, truth, label):
    """Assert the accuracy bound for one array; records the measured numbers."""
    got, ref, fma, truth = (np.asarray(x, dtype=np.float64) for x in (got, ref, fma, truth))
    e_ref = max(_rel(ref, truth), _rel(fma, truth))
    e_gpu = _rel(got, truth)
    d = _rel(got, ref)
    tol = max(RTOL_FLOOR, ACCURACY_FACTOR * e_ref)
    REPORT.append({"case": label, "e_ref": e_ref, "e_gpu": e_gpu, "gpu_vs_ref": d, "tol": tol})
    assert e_gpu <= tol and d <= tol, f"{label}: rel(gpu, truth) {e_gpu:.3e}, rel(gpu, ref) {d:.3e} > tol {tol:.3e} (e_ref {e_ref:.3e})"


def _check_fields(cfg, got, ref, fma, truth, label, skip=()):
    fr, ff, ft = _fields(cfg, ref), _fields(cfg, fma), _fields(cfg, truth)
    for key, exp in fr.items():
        if key in skip:
            continue
        g = got[key][: exp.shape[0]] if key == "output_scale" else got[key]
        check_accuracy(g, exp, ff[key], ft[key], f"{label} {key}")


def _compare(oracle, pdq, cfg, params, y0s, times, fixed_grid=False, check_factor_signs=False, skip=()):
    """Solve a batch on the GPU and every instance with the oracle; compare."""
    B, K = times.shape
    ocfg = _oracle_cfg(oracle, cfg)
    tcs = np.stack([oracle.taylor_coefficients(cfg.vector_field, params[i], y0s[i], times[i, 0], cfg.num_derivatives) for i in range(B)])
    out = pdq.solve(cfg, _dev(params), _dev(tcs), _dev(times), fixed_grid=fixed_grid)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in out.items()}
    n, nblk = cfg.n, cfg.nblk
    for i in range(B):
        r, rf, rt = _oracle_solve(oracle, ocfg, params[i], tcs[i], times[i], fixed_grid)
        assert g["status"][i] == 0, f"instance {i} status {g['status'][i]}"
        assert g["attempts"][i] == r["attempts"], (i, g["attempts"][i], r["attempts"])
        assert g["accepted"][i] == r["accepted"]
        assert np.array_equal(g["num_steps"][i], r["num_steps"])
        assert _rel(g["t"][i], r["t"]) < 1e-14
        got = {k: g[k][i] for k in ("mean", "chol", "output_scale", "filt_mean", "filt_chol", "cond_A", "cond_mean", "term_mean") if k in g}
        _check_fields(cfg, _fields(cfg, got), r, rf, rt, f"{cfg.factorisation}/{cfg.strategy}/{cfg.solver} instance {i}", skip)
        if check_factor_signs:
            # The R-only QR's sign convention ties each column's sign to an
            # element that can be at cancellation level: the strict and FMA
            # oracle builds already disagree on 10% of C1's diagonal signs.
            # Factors are compared up to column signs (positive diagonal).
            def unsigned(ch):
                L = np.asarray(ch, dtype=np.float64).reshape(-1, n, n)
                s = np.sign(np.diagonal(L, axis1=1, axis2=2))
                s[s == 0] = 1.0
                return L * s[:, None, :]

            check_accuracy(unsigned(g["chol"][i]), unsigned(r["chol"]), unsigned(rf["chol"]), unsigned(rt["chol"]), "factor up to column signs")
    return g


def _lv_batch(B, seed, t_lo=5.0, t_hi=15.0, K=20):
    rng = np.random.default_rng(seed)
    params = THETA * np.exp(rng.uniform(np.log(0.8), np.log(1.2), size=(B, 4)))
    y0 = rng.uniform(10, 30, size=(B, 2))
    t_end = rng.uniform(t_lo, t_hi, size=B)
    times = np.stack([np.linspace(0.0, te, K) for te in t_end])
    return params, y0, times


def test_c1_dense_filter_single_solve(oracle, pdq):
    """C1: one LV solve, dense nu=4, ts0, dynamic calibration, filter, 200 save_at."""
    cfg = pdq.Config(factorisation="dense", strategy="filter", solver="dynamic")
    times = np.linspace(0.0, 50.0, 200)[None]
    _compare(oracle, pdq, cfg, THETA[None], np.array([[20.0, 20.0]]), times, check_factor_signs=True)


@pytest.mark.parametrize("lanes", [0, 1, 4, 8, 16, 32, 64])
def test_isotropic_fixedpoint_batch_all_group_sizes(oracle, pdq, lanes):
    """C2 kind: isotropic nu=4, fixed-point smoother, dynamic, uneven lengths.

    lanes 0 (automatic) and 1 run the lane-per-instance kernel, 4..64 the
    cooperative kernel with that group size."""
    cfg = pdq.Config(factorisation="isotropic", strategy="fixedpoint", solver="dynamic", lanes_per_instance=lanes)
    params, y0, times = _lv_batch(12, seed=lanes)
    _compare(oracle, pdq, cfg, params, y0, times)


@pytest.mark.parametrize("fact,lanes", [("dense", 0), ("isotropic", 0), ("isotropic", 16)])
@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
@pytest.mark.parametrize("solver", ["solver", "dynamic", "mle"])
def test_strategy_solver_grid(oracle, pdq, fact, lanes, strategy, solver):
    """isotropic with lanes 0: lane-per-instance kernel; lanes 16: cooperative kernel."""
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7, lanes_per_instance=lanes)
    params, y0, times = _lv_batch(6, seed=7, t_lo=3.0, t_hi=8.0, K=9)
    _compare(oracle, pdq, cfg, params, y0, times)


def test_c3_dense_ts1_van_der_pol(oracle, pdq):
    """C3 kind: van der Pol, dense nu=5, ts1 with the analytic Jacobian."""
    rng = np.random.default_rng(3)
    B = 8
    mu = np.exp(rng.uniform(np.log(1.0), np.log(100.0), size=(B, 1)))
    y0 = np.tile([2.0, 0.0], (B, 1))
    times = np.tile([0.0, 6.3], (B, 1))
    cfg = pdq.Config(factorisation="dense", num_derivatives=5, constraint="ts1", vector_field="van_der_pol", n_params=1, rtol=1e-5, atol=1e-7)
    _compare(oracle, pdq, cfg, mu, y0, times)


@pytest.mark.parametrize("fact,lanes", [("dense", 0), ("isotropic", 0), ("isotropic", 16)])
def test_fixed_grid(oracle, pdq, fact, lanes):
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, strategy="fixedinterval", solver="dynamic", lanes_per_instance=lanes)
    params, y0, _ = _lv_batch(4, seed=11)
    times = np.tile(np.linspace(0.0, 4.0, 41), (4, 1))
    _compare(oracle, pdq, cfg, params, y0, times, fixed_grid=True)


@pytest.mark.skip(reason="this case did not finish within the 120 s test limit on the first blockdiag kernel's first GPU run and its cause was not shown; that kernel is removed and the case is not run on the device again (the current blockdiag kernel is tested by test_gpu_blockdiag.py)")
def test_blockdiag_lotka_volterra(oracle, pdq):
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=4, strategy="filter", solver="dynamic")
    params, y0, times = _lv_batch(5, seed=5)
    _compare(oracle, pdq, cfg, params, y0, times)


@pytest.mark.skip(reason="this case did not finish within the 120 s test limit on the first blockdiag kernel's first GPU run and its cause was not shown; that kernel is removed and the case is not run on the device again (the current blockdiag kernel is tested by test_gpu_blockdiag.py)")
def test_blockdiag_burgers_small_grid(oracle, pdq):
    """C4 kind at a small grid: periodic Burgers, blockdiag nu=2, filter."""
    Ng = 64
    rng = np.random.default_rng(0)
    x = np.arange(Ng) / Ng
    B = 3
    y0 = np.stack([sum(rng.normal(0, 1 / k) * np.sin(2 * np.pi * k * x) for k in range(1, 9)) for _ in range(B)])
    params = np.tile([0.01, 1.0 / Ng], (B, 1))
    times = np.tile([0.0, 0.2], (B, 1))
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=2, ode_dim=Ng, vector_field="burgers", n_params=2, rtol=1e-4, atol=1e-6)
    _compare(oracle, pdq, cfg, params, y0, times)


def test_blockdiag_limits_are_refused(pdq):
    """The C ABI refuses block-diagonal shapes beyond the kernel's limits before any launch."""
    tcs = np.zeros((1, 3, 600))
    for kw, match in [
        (dict(ode_dim=600, vector_field="burgers", n_params=2, num_derivatives=2), "ode_dim"),
        (dict(ode_dim=2, num_derivatives=6), "num_derivatives"),
    ]:
        cfg = pdq.Config(factorisation="blockdiag", strategy="filter", solver="dynamic", **kw)
        p = np.zeros((1, cfg.n_params))
        tc = np.zeros((1, cfg.N, cfg.ode_dim))
        with pytest.raises(pdq.PdqError, match=match):
            pdq.solve(cfg, _dev(p), _dev(tc), _dev(np.array([[0.0, 1.0]])))


def test_oversized_lane_groups_are_refused_and_leave_no_hip_error(oracle, pdq):
    """lanes_per_instance = 4 puts 16 dense nu = 4 instances into one workgroup, more LDS than the
    device allows: the C ABI refuses it as unsupported (no launch), clears the runtime's last error,
    and the next device work (a torch allocation, the same solve with 16 lanes) runs normally."""
    params, y0s, times = _lv_batch(2, seed=3, t_lo=3.0, t_hi=5.0, K=6)
    cfg = pdq.Config(factorisation="dense", num_derivatives=4, strategy="filter", solver="dynamic", lanes_per_instance=4)
    tcs = np.stack([oracle.taylor_coefficients("lotka_volterra", params[i], y0s[i], 0.0, 4) for i in range(2)])
    print("refusal case: dense nu=4 with 4 lanes per instance", flush=True)
    with pytest.raises(pdq.PdqError, match="LDS"):
        pdq.solve(cfg, _dev(params), _dev(tcs), _dev(times))
    z = torch.zeros(8, dtype=torch.float64, device="cuda") + 1.0
    torch.cuda.synchronize()
    assert float(z.sum()) == 8.0
    _compare(oracle, pdq, pdq.Config(factorisation="dense", num_derivatives=4, strategy="filter", solver="dynamic", lanes_per_instance=16), params, y0s, times)


@pytest.mark.parametrize(
    "kw",
    [
        dict(strategy="filter", solver="dynamic", clip_dt=True),
        dict(strategy="filter", solver="dynamic", error_norm="rms_then_scale"),
        dict(strategy="fixedpoint", solver="dynamic", control="proportional_integral"),
    ],
)
def test_lane_kernel_configuration_switches(oracle, pdq, kw):
    """clip_dt, the rms-then-scale error norm and the PI controller on the lane kernel (C2 shape)."""
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=4, **kw)
    params, y0, times = _lv_batch(8, seed=21)
    _compare(oracle, pdq, cfg, params, y0, times)


def test_lane_kernel_refuses_other_shapes(pdq):
    cfg = pdq.Config(factorisation="dense", lanes_per_instance=1)
    params, y0, times = _lv_batch(1, seed=5)
    with pytest.raises(pdq.PdqError, match="lane-per-instance"):
        pdq.solve(cfg, _dev(params), _dev(np.zeros((1, 5, 2))), _dev(times))


def test_taylor_coefficients_match_oracle(oracle, pdq):
    params, y0, _ = _lv_batch(5, seed=2)
    got = pdq.taylor_coefficients("lotka_volterra", _dev(params), _dev(y0), 0.0, 5).cpu().numpy()
    for i in range(5):
        exp = oracle.taylor_coefficients("lotka_volterra", params[i], y0[i], 0.0, 5)
        assert _rel(got[i], exp) < 1e-14


def test_status_max_attempts(pdq):
    cfg = pdq.Config(rtol=1e-300, atol=1e-300, max_attempts=4)
    params, y0, times = _lv_batch(2, seed=1)
    from oracle import oracle as O

    tcs = np.stack([O.taylor_coefficients("lotka_volterra", params[i], y0[i], 0.0, 4) for i in range(2)])
    out = pdq.solve(cfg, _dev(params), _dev(tcs), _dev(times))
    assert out["status"].cpu().numpy().tolist() == [1, 1]


@pytest.mark.parametrize("fact,strategy,lanes", [("isotropic", "fixedpoint", 0), ("isotropic", "filter", 16), ("dense", "filter", 0)])
def test_re_linearize_after_calibration(oracle, pdq, fact, strategy, lanes):
    """solver_dynamic(re_linearize_after_calibration=True) (solvers.py:576-582): both kernels."""
    cfg = pdq.Config(factorisation=fact, num_derivatives=4, strategy=strategy, solver="dynamic", lanes_per_instance=lanes, re_linearize_after_calibration=True)
    params, y0s, times = _lv_batch(4, seed=31, t_lo=5.0, t_hi=10.0, K=15)
    _compare(oracle, pdq, cfg, params, y0s, times)


@pytest.mark.parametrize(
    "fact,lanes,solver,kw",
    [
        ("isotropic", 0, "dynamic", dict(error_estimator="state_std")),
        ("isotropic", 0, "solver", dict(error_estimator="state_std", error_derivative_idx=2, error_re_linearize=True)),
        ("isotropic", 16, "mle", dict(error_re_linearize=True, error_per_unit_step=True)),
        ("dense", 0, "solver", dict(error_re_linearize=True)),
        ("dense", 0, "dynamic", dict(error_per_unit_step=True)),
    ],
)
def test_error_estimator_variants(oracle, pdq, fact, lanes, solver, kw):
    """error_state_std (solvers.py:999-1098; lane kernel) and the estimators' options
    re_linearize_before_error / error_per_unit_step (both kernels)."""
    cfg = pdq.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7, lanes_per_instance=lanes, **kw)
    params, y0s, times = _lv_batch(4, seed=41, t_lo=5.0, t_hi=10.0, K=12)
    _compare(oracle, pdq, cfg, params, y0s, times)


@pytest.mark.parametrize(
    "strategy,solver,kw",
    [
        ("fixedpoint", "dynamic", {}),
        ("filter", "solver", {}),
        ("fixedinterval", "mle", {}),
        ("fixedpoint", "dynamic", dict(error_estimator="state_std", error_derivative_idx=1)),
        ("fixedpoint", "dynamic", dict(re_linearize_after_calibration=True)),
        ("filter", "solver", dict(error_re_linearize=True)),
    ],
)
@pytest.mark.parametrize("lanes", [0, 16])
def test_isotropic_ts1(oracle, pdq, strategy, solver, kw, lanes):
    """Isotropic ts1 with the exact trace (ssm_impl_isotropic.py:321-355, jacobians.py:100-106) on the
    lane kernel (lanes 0, the default) and the cooperative kernel (16 lanes per instance), against the
    oracle pinned by the dense == isotropic equivalence."""
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=4, strategy=strategy, solver=solver, constraint="ts1", rtol=1e-5, atol=1e-7,
                     lanes_per_instance=lanes, **kw)
    params, y0s, times = _lv_batch(6, seed=43, t_lo=5.0, t_hi=10.0, K=12)
    _compare(oracle, pdq, cfg, params, y0s, times)


def test_isotropic_ts1_van_der_pol(oracle, pdq):
    rng = np.random.default_rng(44)
    mu = np.exp(rng.uniform(np.log(1.0), np.log(10.0), size=(6, 1)))
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, strategy="filter", solver="dynamic", constraint="ts1",
                     vector_field="van_der_pol", n_params=1, rtol=1e-5, atol=1e-7)
    _compare(oracle, pdq, cfg, mu, np.tile([2.0, 0.0], (6, 1)), np.tile(np.linspace(0.0, 6.3, 9), (6, 1)))


@pytest.mark.parametrize("nu,strategy", [(2, "fixedpoint"), (6, "fixedinterval")])
def test_isotropic_ts1_shapes_without_a_lane_kernel(oracle, pdq, nu, strategy):
    """Isotropic ts1 on shapes only the cooperative kernel runs (nu = 2 and 6, d = 2)."""
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=nu, strategy=strategy, solver="dynamic", constraint="ts1", rtol=1e-5, atol=1e-7)
    params, y0s, times = _lv_batch(4, seed=53, t_lo=5.0, t_hi=10.0, K=9)
    _compare(oracle, pdq, cfg, params, y0s, times)


@pytest.mark.parametrize("mode", ["monte_carlo_rev", "monte_carlo_fwd"])
@pytest.mark.parametrize(
    "lanes,strategy,kw",
    [
        (0, "fixedpoint", {}),
        (16, "fixedpoint", {}),
        (0, "filter", dict(clip_dt=True)),
        (8, "fixedinterval", dict(re_linearize_after_calibration=True)),
    ],
)
def test_isotropic_ts1_monte_carlo_jacobian(oracle, pdq, mode, lanes, strategy, kw):
    """Isotropic ts1 with the Hutchinson trace of jacobian_monte_carlo_rev / _fwd (jacobians.py:
    151-169, 230-248) on the lane kernel (lanes 0) and the cooperative kernel: the same probes as
    the oracle, one set per linearisation, a rejected attempt's retry reusing its set."""
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=4, strategy=strategy, solver="dynamic", constraint="ts1",
                     rtol=1e-5, atol=1e-7, lanes_per_instance=lanes, jacobian=mode, jac_probes=6, jac_seed=9, **kw)
    params, y0s, times = _lv_batch(6, seed=45, t_lo=5.0, t_hi=10.0, K=12)
    g = _compare(oracle, pdq, cfg, params, y0s, times)
    assert np.all(g["attempts"] > g["accepted"])
    _compare(oracle, pdq, cfg, params, y0s, times, fixed_grid=True)


def test_ts1_lml_gradients_are_refused(pdq):
    """LML gradients need ts0 (the forward-mode tangents share the covariance recursion); ts1 values run."""
    params, y0s, times = _lv_batch(1, seed=1, K=5)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, strategy="fixedpoint", constraint="ts1")
    grid = np.linspace(0.0, 2.0, 11)
    with pytest.raises(pdq.PdqError, match="ts0"):
        pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0s), _dev(grid), 5, _dev(np.ones((1, 3, 2))), _dev(np.ones((1, 3))), grad=True)


def _matern_drift(length_scale, N, d):
    """prior_matern's drift (ssm_impl_api.py:443-477), pinned in tests/test_oracle_prior.py."""
    from math import comb

    z = np.sqrt(2 * (N - 0.5)) / length_scale
    return np.concatenate([-comb(N, i) * z ** (N - i) * np.eye(d) for i in range(N)], axis=1)


@pytest.mark.parametrize(
    "nu,prior,strategy,solver,lanes",
    [
        (3, "zero", "filter", "dynamic", 0),
        (4, "matern", "filter", "dynamic", 0),
        (3, "ioup", "fixedpoint", "dynamic", 0),
        (3, "matern", "fixedinterval", "solver", 16),
        (4, "ioup", "filter", "mle", 64),
        (7, "matern", "filter", "dynamic", 0),
    ],
)
def test_exponential_prior(oracle, pdq, nu, prior, strategy, solver, lanes):
    """The dense exponential prior (ssm_impl_dense.py:392-629): every transition is
    exp_gram_cholesky(pade_and_legendre_9), computed per step on the device, against the oracle
    pinned by the matrix-fraction baseline and the zero-drift == IWP equivalence."""
    d = 2
    N, n = nu + 1, (nu + 1) * d
    if prior == "zero":
        drift = np.zeros((d, n))
    elif prior == "matern":
        drift = _matern_drift(8.0, N, d)
    else:
        drift = np.zeros((d, n))
        drift[:, nu * d :] = np.array([[-0.3, 0.1], [0.05, -0.2]])
    cfg = pdq.Config(factorisation="dense", num_derivatives=nu, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7,
                     lanes_per_instance=lanes, prior="exponential", prior_drift=_dev(drift))
    params, y0s, times = _lv_batch(6, seed=45, t_lo=4.0, t_hi=8.0, K=10)
    _compare(oracle, pdq, cfg, params, y0s, times)


def test_exponential_prior_refusals(pdq):
    """Dense only (as the reference), (nu + 1) d <= 16 and a drift pointer.  (The LML runs them:
    tests/test_gpu_lml.py.)"""
    params, y0s, times = _lv_batch(1, seed=1, K=5)
    tc = pdq.taylor_coefficients("lotka_volterra", _dev(params), _dev(y0s), 0.0, 3)
    drift = _dev(np.zeros((2, 8)))
    for kw, match in [
        (dict(factorisation="isotropic", prior_drift=drift), "dense factorisation"),
        (dict(factorisation="dense"), "prior_drift"),
    ]:
        cfg = pdq.Config(num_derivatives=3, prior="exponential", **kw)
        with pytest.raises(pdq.PdqError, match=match):
            pdq.solve(cfg, _dev(params), tc, _dev(times))
    # (nu + 1) d = 18
    lp, ly = _dev(np.full((1, 1), 1.5)), _dev(np.full((1, 3), 0.3))
    cfg = pdq.Config(factorisation="dense", num_derivatives=5, ode_dim=3, vector_field="logistic", n_params=1,
                     prior="exponential", prior_drift=_dev(np.zeros((3, 18))))
    with pytest.raises(pdq.PdqError, match="<= 16"):
        pdq.solve(cfg, lp, pdq.taylor_coefficients("logistic", lp, ly, 0.0, 5), _dev(times))


@pytest.mark.parametrize(
    "fact,solver,lanes,kw",
    [
        ("dense", "dynamic", 0, dict(error_derivative_idx=0)),
        ("dense", "dynamic", 0, dict(error_derivative_idx=2, re_linearize_after_calibration=True)),
        ("dense", "solver", 0, dict(error_derivative_idx=0, error_re_linearize=True)),
        ("dense", "mle", 0, dict(error_derivative_idx=3, error_per_unit_step=True, error_norm="rms_then_scale")),
        ("isotropic", "dynamic", 16, dict(error_derivative_idx=2)),
    ],
)
def test_error_state_std_on_the_cooperative_kernel(oracle, pdq, fact, solver, lanes, kw):
    """error_state_std on the cooperative kernel: dense (d calibrated stds of derivative idx) and the
    isotropic group path, against the oracle.  Derivative 1 is left out: the ODE constraint pins it,
    so its posterior std is rounding noise and the step sequence depends on the rounding."""
    cfg = pdq.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7,
                     error_estimator="state_std", lanes_per_instance=lanes, **kw)
    params, y0s, times = _lv_batch(4, seed=47, t_lo=5.0, t_hi=10.0, K=12)
    _compare(oracle, pdq, cfg, params, y0s, times)


@pytest.mark.parametrize("fact,lanes", [("dense", 0), ("isotropic", 0), ("isotropic", 16)])
def test_init_std_rows(oracle, pdq, fact, lanes):
    """A std per covariance row at t0 (element-wise is_exact, diffuse_derivatives): the lane kernel's
    option instantiation (isotropic, lanes 0) and the cooperative kernel, against the oracle."""
    rows = np.array([0.0, 0.0, 1e-3, 1e-3, 0.5, 0.5, 1.0, 1.0, 1.0, 1.0]) if fact == "dense" else np.array([0.0, 0.0, 1e-3, 0.5, 1.0])
    cfg = pdq.Config(factorisation=fact, num_derivatives=4, strategy="fixedpoint", solver="dynamic", rtol=1e-5, atol=1e-7,
                     lanes_per_instance=lanes, init_std_rows=_dev(rows))
    params, y0s, times = _lv_batch(4, seed=57, t_lo=5.0, t_hi=10.0, K=12)
    _compare(oracle, pdq, cfg, params, y0s, times)
