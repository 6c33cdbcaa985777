"""CPU-side checks of the C ABI and the host logic (no compute calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "pdq.h")).read()
    return sorted(set(re.findall(r"\b(pdq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from probdiffeq_amd import capi

    lib = capi.lib()
    declared = _header_symbols()
    assert "pdq_solve_adaptive_save_at" in declared
    for name in declared:
        assert hasattr(lib, name), name
    for name in capi.EXPORTED:
        assert name in declared, name


def test_config_struct_matches_header_layout():
    """The ctypes mirror of pdq_config has the C layout (12 ints, 12 doubles, 8 ints, two pointers,
    one int padded to the struct's 8-byte alignment, one pointer, five ints padded)."""
    from probdiffeq_amd import capi

    assert ctypes.sizeof(capi._Config) == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8 + 8 + 8 + 24
    assert capi._Config.lml_gradient.offset == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8 + 8 + 8
    assert capi._Config.lml_solve.offset == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8 + 8 + 8 + 4
    assert capi._Config.jac_seed.offset == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8 + 8 + 8 + 16
    assert capi._Config.lml_tcoeff_index.offset == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8
    assert capi._Config.init_std_rows.offset == 12 * 4 + 12 * 8 + 8 * 4 + 2 * 8 + 8
    assert ctypes.sizeof(capi._Solution) == 17 * 8


def test_sizes_match_layouts():
    from probdiffeq_amd import capi

    lib = capi.lib()
    for fact, n_chol in [("dense", 100), ("isotropic", 25), ("blockdiag", 2 * 25)]:
        cfg = capi.Config(factorisation=fact, num_derivatives=4, ode_dim=2)
        c = cfg.c()
        assert lib.pdq_mean_size(ctypes.byref(c)) == 10 == cfg.mean_size
        assert lib.pdq_chol_size(ctypes.byref(c)) == n_chol == cfg.chol_size


def test_no_device_is_an_error_not_a_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from probdiffeq_amd import capi

    assert capi.lib().pdq_init() == capi.ERR_NO_DEVICE
    with pytest.raises(capi.PdqError, match="no HIP device"):
        capi.require_device()


def test_validation_errors_are_reported():
    """Argument checks run before any device work (controllers and shapes)."""
    from probdiffeq_amd import capi

    lib = capi.lib()
    cfg = capi.Config(factorisation="isotropic", prior="exponential")
    c = cfg.c()
    sol = capi._Solution()
    rc = lib.pdq_solve_adaptive_save_at(ctypes.byref(c), 1, None, None, None, 10, ctypes.byref(sol), None)
    assert rc == 2  # PDQ_ERR_UNSUPPORTED
    assert "dense factorisation" in capi.last_error()
    cfg = capi.Config(num_derivatives=12)
    c = cfg.c()
    rc = lib.pdq_solve_adaptive_save_at(ctypes.byref(c), 1, None, None, None, 10, ctypes.byref(sol), None)
    assert rc == 2 and "num_derivatives" in capi.last_error()
    rc = lib.pdq_solve_adaptive_save_at(ctypes.byref(capi.Config().c()), 1, None, None, None, 1, ctypes.byref(sol), None)
    assert rc == 1 and "two time points" in capi.last_error()


def test_api_objects_mirror_the_reference():
    """Configuration objects and their reference error behaviour (no device needed)."""
    from probdiffeq_amd import probdiffeq as pdq

    assert pdq.strategy_smoother_fixedpoint().is_suitable_for_save_every_step is False
    assert pdq.strategy_smoother_fixedinterval().is_suitable_for_save_at is False
    ssm = pdq.state_space_model_isotropic()
    with pytest.raises(TypeError):
        ssm.constraint_ode_ts0(lambda y, t: y)
    with pytest.raises(NotImplementedError):
        pdq.ode(lambda y, *, t: y)
    assert pdq.solver_dynamic(strategy=pdq.strategy_filter(), constraint=None, re_linearize_after_calibration=True).re_linearize_after_calibration
    # ts1 takes ode()'s default handler (jacobian_monte_carlo_rev(seed=1, num_probes=10)) or the
    # one the ODE names, for every factorisation
    lv = pdq.JetOde("lotka_volterra", None, 2)  # registry entry; its parameters stay unset here
    assert pdq._jacobian_handler(lv.jacobian) == ("monte_carlo_rev", 10, 1)
    assert pdq._jacobian_handler(pdq.jacobian_monte_carlo_fwd(seed=7, num_probes=3)) == ("monte_carlo_fwd", 3, 7)
    assert pdq._jacobian_handler(pdq.jacobian_materialize())[0] == "exact"
    for model in (ssm, pdq.state_space_model_dense(), pdq.state_space_model_blockdiag()):
        assert model.constraint_ode_ts1(lv).kind == "ts1"
        assert model.constraint_ode_ts1(pdq.ode(lv, jacobian=pdq.jacobian_materialize())).kind == "ts1"
    with pytest.raises(NotImplementedError, match="probes"):
        ssm.constraint_ode_ts1(pdq.ode(lv, jacobian=pdq.jacobian_monte_carlo_rev(num_probes=65)))
    with pytest.raises(TypeError):
        ssm.constraint_ode_ts1(pdq.ode(lv, jacobian=object()))


def test_fixed_grid_warns_for_fixedpoint_smoother():
    """test_solve_fixed_grid.py:57-71."""
    from probdiffeq_amd import ivpsolve
    from probdiffeq_amd import probdiffeq as pdq

    slv = pdq.solver(strategy=pdq.strategy_smoother_fixedpoint(), constraint=None)
    with pytest.warns(UserWarning, match="fixed-point"):
        ivpsolve.solve_fixed_grid(solver=slv)


def test_tuning_table_accepts_kernel_choices_and_rejects_others():
    """pdq_set_tuned_lanes needs no device: 1 (lane kernel) and 4..64 are recorded, 3 is refused."""
    from probdiffeq_amd import capi

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4)
    c = cfg.c()
    assert capi.lib().pdq_set_tuned_lanes(ctypes.byref(c), 100, 0, 1) == 0
    assert capi.lib().pdq_tuned_lanes(ctypes.byref(c), 100) == 1
    assert capi.lib().pdq_set_tuned_lanes(ctypes.byref(c), 100, 0, 32) == 0
    assert capi.lib().pdq_tuned_lanes(ctypes.byref(c), 128) == 32  # same batch-size bucket
    assert capi.lib().pdq_tuned_lanes(ctypes.byref(c), 129) == 0
    assert capi.lib().pdq_set_tuned_lanes(ctypes.byref(c), 100, 0, 3) == 1  # PDQ_ERR_INVALID
    assert "lane kernel" in capi.last_error()


def test_lml_terminal_values_checks_the_std_shape_like_the_reference():
    """estimators_and_losses.py:24-42: a std container of the wrong shape is a ValueError (raised
    before any device call)."""
    import torch

    from probdiffeq_amd import probdiffeq as pdq

    marg = pdq.Normal(torch.zeros(5, 2, dtype=torch.float64), torch.eye(5, dtype=torch.float64), "isotropic", 5, 2)
    loss = pdq.loss_lml_terminal_values()
    with pytest.raises(ValueError, match="standard deviation container"):
        loss(torch.zeros(2, dtype=torch.float64), marginals=marg, std=torch.ones(2, dtype=torch.float64))


def test_exponential_prior_drifts_match_the_reference_closed_forms():
    """The mirror's drift matrices (the Jacobian of the prior's linear ODE, ssm_impl_dense.py:587-596):
    Matern (test_matern.py:7-27: f = -l^3 u - 3 l^2 du - 3 l ddu, l = sqrt(5) / length_scale),
    IOUP (test_ornstein_uhlenbeck_integrated.py: f = M u^(nu)) and an arbitrary-order linear ODE."""
    import numpy as np
    import torch

    from probdiffeq_amd import probdiffeq as pdq

    tc = torch.ones(3, 4, dtype=torch.float64)
    ls = 0.567
    ell = np.sqrt(2 * 2.5) / ls
    expected = np.concatenate([-(ell**3) * np.eye(4), -3 * ell**2 * np.eye(4), -3 * ell * np.eye(4)], axis=1)
    assert np.allclose(pdq._linear_drift(pdq._matern_ode(ls, 3), tc).numpy(), expected, rtol=1e-15, atol=0)
    M = torch.as_tensor(np.random.default_rng(1).normal(size=(4, 4)))
    J = pdq._linear_drift(pdq._ioup_ode(lambda x: M @ x, 3), tc).numpy()
    assert np.array_equal(J[:, :8], np.zeros((4, 8))) and np.allclose(J[:, 8:], M.numpy(), rtol=1e-15, atol=0)
    ode = pdq.ode_autonomous_order_arbitrary(lambda u, du, ddu, /: M @ ddu - 2.0 * u, num_tcoeffs_in_args=3)
    J = pdq._linear_drift(ode, tc).numpy()
    assert np.allclose(J, np.concatenate([-2.0 * np.eye(4), np.zeros((4, 4)), M.numpy()], axis=1), rtol=1e-15, atol=0)
    with pytest.raises(NotImplementedError, match="exponential priors"):
        pdq.state_space_model_isotropic().prior_exponential(ode, [tc[0]] * 3)


def test_system_matrices_and_preconditioner_match_the_oracle(oracle):
    """utilities.system_matrices_1d_iwp / preconditioner_taylor of the mirror (the library's host
    matrices; no device) against the oracle, which is pinned to mpmath and the reference's KAT."""
    import numpy as np

    from probdiffeq_amd import probdiffeq as pdq

    for nu in range(0, 8):
        A, Q = pdq.system_matrices_1d_iwp(nu)
        Ao, Qo = oracle.system_matrices(nu)
        assert np.array_equal(A.numpy(), Ao)
        assert np.allclose(Q.numpy(), Qo, rtol=0, atol=1e-15 * np.max(np.abs(Qo)))
    for dt in (1.234, -1.234, 0.01):
        p, pinv = pdq.preconditioner_taylor(4)(dt)
        assert np.allclose(p.numpy(), [dt**4 / 24, dt**3 / 6, dt**2 / 2, dt, 1.0], rtol=1e-15, atol=0)
        assert np.allclose((p * pinv).numpy(), 1.0, rtol=1e-15, atol=0)
    assert pdq.state_space_model_dense().constraint_ode_ts1(pdq.JetOde("lotka_volterra", None, 2), taylor_point=pdq.taylor_point_prior()).kind == "ts1"



def test_device_read_config_fields_must_be_device_tensors():
    """init_std_rows, prior_drift and dt0_per_instance are read by the kernels on the GPU: the
    binding refuses host memory for them (numpy arrays, CPU tensors) and wrong shapes or dtypes,
    before any call into the library; the CPU build of the kernels takes numpy (host=True)."""
    import numpy as np
    import torch

    from probdiffeq_amd import capi

    for field, value in [
        ("init_std_rows", np.zeros(10)),
        ("init_std_rows", torch.zeros(10, dtype=torch.float64)),
        ("prior_drift", np.zeros((2, 10))),
        ("dt0_per_instance", torch.full((3,), 0.1, dtype=torch.float64)),
    ]:
        cfg = capi.Config(factorisation="dense", **{field: value})
        with pytest.raises(capi.PdqError, match="device"):
            cfg.c()
        cfg.c(host=True) if isinstance(value, np.ndarray) else None
    with pytest.raises(ValueError, match="shape"):
        capi.Config(factorisation="dense", init_std_rows=np.zeros(7)).c(host=True)
    with pytest.raises(ValueError, match="shape"):
        capi.Config(factorisation="dense", prior_drift=np.zeros((10, 2))).c(host=True)
    with pytest.raises(ValueError, match="dimensions"):
        capi.Config(dt0_per_instance=np.zeros((2, 2))).c(host=True)


def test_solve_wrappers_refuse_host_tensors():
    import torch

    from probdiffeq_amd import capi

    cfg = capi.Config()
    x = torch.zeros(2, 4, dtype=torch.float64)
    with pytest.raises(capi.PdqError, match="device"):
        capi.solve(cfg, x, torch.zeros(2, 5, 2, dtype=torch.float64), torch.zeros(2, 3, dtype=torch.float64))
    with pytest.raises(capi.PdqError, match="device"):
        capi.lml_value_and_grad(cfg, x, torch.zeros(2, 2), torch.zeros(11), 5, torch.zeros(2, 3, 2), torch.zeros(2, 3))


@pytest.mark.parametrize(
    "vf,d,n_params,num,match",
    [
        (0, 3, 4, 4, "ode_dim == 2"),  # lotka_volterra with d != 2
        (0, 2, 3, 4, "4 params"),  # lotka_volterra with too few params
        (1, 2, 0, 4, "1 param"),  # van_der_pol without mu
        (3, 4, 2, 2, "d params"),  # linear with n_params < d
        (9, 2, 4, 4, "unknown vector field"),
        (0, 2, 4, -1, "invalid sizes"),
        (0, 2, 4, 2**30, "num must not exceed 7"),  # (num + 1) d would overflow an int
        (2, 2**21, 2, 1, "2^20"),
    ],
)
def test_taylor_coefficients_checks_its_arguments(vf, d, n_params, num, match):
    """pdq_taylor_coefficients applies the vector-field registry's requirements and bounds num and d
    before any device work (its kernel indexes params and the coefficient rows by them)."""
    from probdiffeq_amd import capi

    lib = capi.lib()
    rc = lib.pdq_taylor_coefficients(vf, d, n_params, 1, None, None, ctypes.c_double(0.0), num, None, None)
    assert rc in (1, 2), rc
    assert match in capi.last_error()


def test_lml_gradient_mode_is_checked_before_device_work():
    """pdq_config.lml_gradient: the binding accepts "forward" and "reverse"; the C ABI refuses an
    unknown mode and the shapes the reverse adjoint does not cover before touching the device."""
    from probdiffeq_amd import capi

    with pytest.raises(ValueError, match="lml_gradient"):
        capi.Config(lml_gradient="backward").c()
    lib = capi.lib()
    fake = ctypes.c_void_p(256)  # never dereferenced: the checks fail first

    def call(c):
        return lib.pdq_lml_value_and_grad(ctypes.byref(c), 1, fake, fake, fake, 11, 5, fake, fake, 1, 1, fake, fake, None)

    base = dict(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver")
    c = capi.Config(**base).c()
    c.lml_gradient = 7
    assert call(c) == 1 and "unknown lml_gradient" in capi.last_error()
    for kw, match in [
        (dict(factorisation="dense"), "reverse adjoint"),
        (dict(num_derivatives=5), "reverse adjoint"),
        (dict(lml_tcoeff_index=2), "reverse adjoint"),
        (dict(lanes_per_instance=8), "one lane per instance"),
    ]:
        cfg = dict(base, lml_gradient="reverse")
        cfg.update(kw)
        assert call(capi.Config(**cfg).c()) == 2, kw
        assert match in capi.last_error(), (kw, capi.last_error())


def test_jacobian_handler_is_checked_before_device_work():
    """pdq_config.jacobian / jac_probes: unknown handlers, probe counts outside 1..64 and the
    re-linearisation before the error estimate (which draws from the error estimator's own key)
    are refused before the device is touched."""
    from probdiffeq_amd import capi

    with pytest.raises(ValueError, match="jacobian"):
        capi.Config(jacobian="hutchinson").c()
    lib = capi.lib()
    sol = capi._Solution()

    def call(c):
        return lib.pdq_solve_adaptive_save_at(ctypes.byref(c), 1, None, None, None, 10, ctypes.byref(sol), None)

    base = dict(factorisation="isotropic", num_derivatives=3, constraint="ts1", jacobian="monte_carlo_rev")
    c = capi.Config(**base).c()
    c.jacobian = 5
    assert call(c) == 1 and "unknown jacobian" in capi.last_error()
    for probes in (0, 65):
        assert call(capi.Config(**base, jac_probes=probes).c()) == 1
        assert "jac_probes" in capi.last_error()
    for fact in ("isotropic", "blockdiag"):
        cfg = dict(base, factorisation=fact, error_re_linearize=True, solver="solver")
        assert call(capi.Config(**cfg).c()) == 2 and "re_linearize_before_error" in capi.last_error()
    # the dense factorisation materialises the Jacobian: the option combination passes the checks
    # (and then stops at the missing device or the null inputs)
    cfg = dict(base, factorisation="dense", error_re_linearize=True, solver="solver")
    assert call(capi.Config(**cfg).c()) not in (0, 2) or "re_linearize" not in capi.last_error()
