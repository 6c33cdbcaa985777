"""GPU: the block-diagonal factorisation (ssm_impl_blockdiag.py) against the oracle.

The device kernel (probdiffeq_amd/csrc/pdq_bd.h) runs one lane per block and one
lane group per instance.  These tests cover both group shapes (a wavefront
shared by 64 / G instances for d <= 64, one workgroup per instance for
d > 64), every strategy and solver, ts1 with the exact Jacobian diagonal, the
error-estimator and controller options, fixed grids, and the C4 configuration
(periodic Burgers on 512 points, nu = 2, rtol 1e-4, atol 1e-6) at a short
horizon against the oracle and at its full horizon through size-independent
properties and a two-instance oracle comparison.
"""

import numpy as np
import pytest
import torch

import test_gpu_parity as T

pytestmark = pytest.mark.gpu

_dev = T._dev


@pytest.fixture(scope="module")
def pdq():
    from probdiffeq_amd import capi

    capi.require_device()
    capi.lib()
    return capi


def _burgers(Ng, B, seed=0, nmodes=8):
    """C4's initial condition on Ng points: sum of nmodes sinusoids with N(0, 1/k^2) amplitudes
    (phases 0), periodic grid x_j = j / Ng; parameters (viscosity 0.01, dx)."""
    rng = np.random.default_rng(seed)
    x = np.arange(Ng) / Ng
    y0 = np.stack([sum(rng.normal(0, 1 / k) * np.sin(2 * np.pi * k * x) for k in range(1, nmodes + 1)) for _ in range(B)])
    return np.tile([0.01, 1.0 / Ng], (B, 1)), y0


def _c4_cfg(pdq, **kw):
    base = dict(factorisation="blockdiag", num_derivatives=2, ode_dim=512, vector_field="burgers", n_params=2,
                strategy="filter", solver="dynamic", rtol=1e-4, atol=1e-6)
    base.update(kw)
    return pdq.Config(**base)


def test_c4_configuration_short_horizon(oracle, pdq):
    """C4 (512 blocks, one 512-lane workgroup per instance) on t in [0, 0.02] against the oracle:
    identical step sequences and the accuracy bound of test_gpu_parity."""
    params, y0 = _burgers(512, 2, seed=0)
    T._compare(oracle, pdq, _c4_cfg(pdq), params, y0, np.tile([0.0, 0.01, 0.02], (2, 1)))


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
def test_multi_wave_group_with_padding_lanes(oracle, pdq, strategy):
    """d = 100: one workgroup of 128 lanes per instance, 28 of them padding.

    With the fixed-point smoother the merged backward gains (cond_A) of this stiff problem are
    not determined by FP64: the strict and FMA builds of the oracle already differ from the long
    double solution by 2e-4 to 0.7 relative on them (the merged interval maps x(t1) back to x(t0)
    through many dissipative steps), so they are left out of that comparison; the smoothed means
    and covariances, which apply them, are compared."""
    params, y0 = _burgers(100, 3, seed=3)
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=2, ode_dim=100, vector_field="burgers", n_params=2,
                     strategy=strategy, rtol=1e-4, atol=1e-6)
    skip = ("cond_A",) if strategy == "fixedpoint" else ()
    T._compare(oracle, pdq, cfg, params, y0, np.tile([0.0, 0.04, 0.1], (3, 1)), skip=skip)


@pytest.mark.parametrize("strategy", ["filter", "fixedpoint", "fixedinterval"])
@pytest.mark.parametrize("solver", ["solver", "dynamic", "mle"])
def test_strategy_solver_grid(oracle, pdq, strategy, solver):
    """Lotka-Volterra, d = 2 (32 instances per wavefront), nu = 3, uneven horizons."""
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=3, strategy=strategy, solver=solver, rtol=1e-5, atol=1e-7)
    params, y0, times = T._lv_batch(6, seed=41, t_lo=3.0, t_hi=8.0, K=9)
    T._compare(oracle, pdq, cfg, params, y0, times)


@pytest.mark.parametrize(
    "vf,d,strategy",
    [("logistic", 5, "filter"), ("logistic", 5, "fixedpoint"), ("logistic", 37, "fixedinterval"), ("burgers", 16, "filter")],
)
def test_ts1_exact_jacobian_diagonal(oracle, pdq, vf, d, strategy):
    """Blockdiag ts1 (BlockDiagResidual.linearize, ssm_impl_blockdiag.py:147-183, with
    jacobian_materialize's diagonal, jacobians.py:108-113), against the oracle pinned by the
    vmap(dense) == blockdiag equivalence (tests/test_oracle_solves.py)."""
    rng = np.random.default_rng(43)
    B = 4
    if vf == "logistic":
        params = np.full((B, 1), 1.5)
        y0 = rng.uniform(0.1, 0.9, size=(B, d))
        times = np.tile(np.linspace(0.0, 4.0, 6), (B, 1))
    else:
        params, y0 = _burgers(d, B, seed=5, nmodes=3)
        times = np.tile([0.0, 0.05, 0.1], (B, 1))
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=3, ode_dim=d, vector_field=vf, n_params=params.shape[1],
                     constraint="ts1", strategy=strategy, rtol=1e-5, atol=1e-7)
    T._compare(oracle, pdq, cfg, params, y0, times)


@pytest.mark.parametrize(
    "mode,vf,d,strategy,kw",
    [
        ("monte_carlo_rev", "lotka_volterra", 2, "fixedpoint", {}),
        ("monte_carlo_fwd", "lotka_volterra", 2, "filter", dict(re_linearize_after_calibration=True)),
        ("monte_carlo_rev", "burgers", 16, "filter", {}),
        ("monte_carlo_fwd", "burgers", 70, "fixedinterval", {}),
    ],
)
def test_ts1_monte_carlo_jacobian_diagonal(oracle, pdq, mode, vf, d, strategy, kw):
    """Blockdiag ts1 with the Hutchinson diagonal of jacobian_monte_carlo_rev / _fwd (jacobians.py:
    171-192, 250-271): every block's lane draws the oracle's probes (d = 70 spans two wavefronts)."""
    B = 3
    if vf == "burgers":
        params, y0 = _burgers(d, B, seed=6, nmodes=3)
        times = np.tile([0.0, 0.05, 0.1], (B, 1))
    else:
        params, y0, times = T._lv_batch(B, seed=47, t_lo=3.0, t_hi=6.0, K=7)
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=3, ode_dim=d, vector_field=vf, n_params=params.shape[1],
                     constraint="ts1", strategy=strategy, rtol=1e-5, atol=1e-7, jacobian=mode, jac_probes=5, jac_seed=4, **kw)
    T._compare(oracle, pdq, cfg, params, y0, times)


@pytest.mark.parametrize("strategy", ["filter", "fixedinterval"])
def test_fixed_grid(oracle, pdq, strategy):
    params, y0, _ = T._lv_batch(4, seed=47)
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=3, strategy=strategy, solver="dynamic")
    T._compare(oracle, pdq, cfg, params, y0, np.tile(np.linspace(0.0, 4.0, 41), (4, 1)), fixed_grid=True)


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
def test_options(oracle, pdq, solver, kw):
    cfg = pdq.Config(factorisation="blockdiag", num_derivatives=3, strategy="fixedpoint", solver=solver, rtol=1e-5, atol=1e-7, **kw)
    params, y0, times = T._lv_batch(4, seed=53, t_lo=3.0, t_hi=6.0, K=7)
    T._compare(oracle, pdq, cfg, params, y0, times)


def _c4_batch(pdq, B, seed=0):
    params, y0 = _burgers(512, B, seed=seed)
    p, y = _dev(params), _dev(y0)
    tc = pdq.taylor_coefficients("burgers", p, y, 0.0, 2)
    return params, y0, p, tc


@pytest.mark.timeout(300)
def test_c4_full_horizon(oracle, pdq):
    """C4 on t in [0, 1], 64 instances in one launch: every instance ends OK, each instance's result
    is bitwise the one it gets when solved alone (the group shares nothing across instances), and
    two instances take the oracle's step sequence with u(1) within 1e-8 of it (the oracle needs
    about a minute per instance here)."""
    B = 64
    params, y0, p, tc = _c4_batch(pdq, B)
    times = _dev(np.tile([0.0, 1.0], (B, 1)))
    cfg = _c4_cfg(pdq)
    out = pdq.solve(cfg, p, tc, times)
    torch.cuda.synchronize()
    assert int(out["status"].abs().sum()) == 0
    att = out["attempts"].cpu().numpy()
    assert att.min() > 1000
    for i in (0, 37):
        one = pdq.solve(cfg, p[i : i + 1].contiguous(), tc[i : i + 1].contiguous(), times[i : i + 1].contiguous())
        torch.cuda.synchronize()
        for name in ("mean", "chol", "output_scale", "attempts"):
            assert torch.equal(one[name][0], out[name][i]), (i, name)
    ocfg = T._oracle_cfg(oracle, cfg)
    tcs = tc.cpu().numpy()
    ref = oracle.solve_save_at_batch(ocfg, params[:2], tcs[:2], np.tile([0.0, 1.0], (2, 1)), n_threads=2)
    assert ref["status"] == 0
    mean = out["mean"].cpu().numpy()
    for i in range(2):
        assert att[i] == ref["attempts"][i], (i, att[i], ref["attempts"][i])
        u_gpu = mean[i, -1].reshape(512, 3)[:, 0]
        u_ref = ref["term_mean"][i].reshape(512, 3)[:, 0]
        assert T._rel(u_gpu, u_ref) < 1e-8


def test_c4_matches_dop853_at_the_solver_tolerance(pdq):
    """The C4 solution u(1) against a high-accuracy DOP853 solution of the same ODE
    (test_solve_adaptive_save_at.py:211-247's accuracy check at C4's rtol 1e-4)."""
    from scipy.integrate import solve_ivp

    params, y0, p, tc = _c4_batch(pdq, 2, seed=1)
    out = pdq.solve(_c4_cfg(pdq), p, tc, _dev(np.tile([0.0, 1.0], (2, 1))))
    torch.cuda.synchronize()
    dx, visc = 1.0 / 512, 0.01

    def f(_, u):
        up, um = np.roll(u, -1), np.roll(u, 1)
        return -u * ((up - um) / (2 * dx)) + visc * ((up - 2 * u + um) / (dx * dx))

    for i in range(2):
        s = solve_ivp(f, (0.0, 1.0), y0[i], method="DOP853", rtol=1e-10, atol=1e-12)
        u = out["mean"][i, -1].cpu().numpy().reshape(512, 3)[:, 0]
        err = np.max(np.abs(u - s.y[:, -1])) / np.max(np.abs(s.y[:, -1]))
        assert err < 1e-3, err


def test_taylor_coefficients_of_a_large_system(oracle, pdq):
    """pdq_taylor_coefficients beyond 512 coefficients (C4: 3 x 512) forms them in its output rows;
    against the oracle's recursion."""
    params, y0 = _burgers(512, 3, seed=2)
    tc = pdq.taylor_coefficients("burgers", _dev(params), _dev(y0), 0.0, 2).cpu().numpy()
    for i in range(3):
        ref = oracle.taylor_coefficients("burgers", params[i], y0[i], 0.0, 2)
        assert T._rel(tc[i], ref) < 1e-13
