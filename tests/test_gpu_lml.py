"""GPU: marginal log-likelihood value and gradient through the C ABI.

Value: against the oracle's loss_lml_timeseries restatement in FP64 and in
long double, with the accuracy bound of the solve parity tests.  Gradient (forward
mode through the solve and the reverse LML scan): against Richardson
finite differences of the long double oracle, 1e-10 normwise relative.
"""

import numpy as np
import pytest

import lml_cases as LC
import test_gpu_parity as T

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pdq():
    from probdiffeq_amd import capi

    capi.require_device()
    capi.lib()
    return capi


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda()


def _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, average=True, grad=True, fd_spread_tol=False):
    ocfg = T._oracle_cfg(oracle, cfg)
    print(f"lml case: fact={cfg.factorisation} solver={cfg.solver} lanes={cfg.lanes_per_instance} B={params.shape[0]} K={grid.size}", flush=True)
    v, g = pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds), average=average, grad=grad)
    torch.cuda.synchronize()
    v = v.cpu().numpy()
    g = g.cpu().numpy() if grad else None
    for i in range(params.shape[0]):
        vo, vf, vt = LC.oracle_value(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
        T.check_accuracy(v[i], vo, vf, vt, f"lml {cfg.factorisation}/{cfg.solver} instance {i}")
        if grad and fd_spread_tol:
            gr, tol = LC.grad_reference(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
            assert tol < 1e-8
            assert LC.rel(g[i], gr) < tol, (i, g[i], gr, tol)
        elif grad:
            gr = LC.richardson_grad(oracle, ocfg, params[i], y0[i], grid, every, data[i], stds[i], average)
            assert LC.rel(g[i], gr) < LC.GRAD_RTOL, (i, g[i], gr)
    return v, g


@pytest.mark.parametrize("lanes", [0, 1, 8, 64])
def test_lml_isotropic_batch(oracle, pdq, lanes):
    params, y0, grid, every, data, stds = LC.lml_inputs(6, seed=lanes)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver", lanes_per_instance=lanes)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds)


@pytest.mark.parametrize(
    "fact,solver,per_coordinate_std,average",
    [("isotropic", "dynamic", False, False), ("isotropic", "dynamic", False, True), ("dense", "solver", True, True), ("dense", "dynamic", False, True)],
)
def test_lml_variants(oracle, pdq, fact, solver, per_coordinate_std, average):
    params, y0, grid, every, data, stds = LC.lml_inputs(4, seed=5, t_end=10.0, n_grid=1001, every=10, per_coordinate_std=per_coordinate_std)
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver=solver)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, average=average)


@pytest.mark.parametrize("fact,idx,lanes,grad", [("isotropic", 1, 0, True), ("isotropic", 2, 16, True), ("dense", 2, 0, True), ("dense", 3, 0, False)])
def test_lml_observing_a_derivative(oracle, pdq, fact, idx, lanes, grad):
    """loss_lml_timeseries(tcoeff_index=idx): the data observe Taylor coefficient idx.  Isotropic
    nu = 3, d = 2 with lanes 0 would take the lane LML for idx = 0; idx > 0 runs the cooperative one.
    idx = nu: value only (finite differences of the oracle cannot pin that gradient to 1e-10)."""
    params, y0, grid, every, data, stds = LC.lml_inputs(4, seed=11, t_end=10.0, n_grid=1001, every=10)
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic",
                     lanes_per_instance=lanes, lml_tcoeff_index=idx)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=grad, fd_spread_tol=True)


@pytest.mark.parametrize("fact,lanes", [("isotropic", 0), ("isotropic", 16), ("dense", 0)])
def test_lml_value_with_ts1(oracle, pdq, fact, lanes):
    """The LML value with ts1 (isotropic: the exact trace) runs the cooperative LML; the lane LML is ts0 only."""
    params, y0, grid, every, data, stds = LC.lml_inputs(4, seed=13, t_end=10.0, n_grid=1001, every=10)
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic",
                     constraint="ts1", lanes_per_instance=lanes)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=False)


@pytest.mark.parametrize("lanes", [0, 16])
def test_lml_value_with_ts1_monte_carlo_jacobian(oracle, pdq, lanes):
    """The LML value with isotropic ts1 and ode()'s default handler, jacobian_monte_carlo_rev: on the
    fixed grid step k draws probe set k - 1, as in the oracle."""
    params, y0, grid, every, data, stds = LC.lml_inputs(4, seed=15, t_end=10.0, n_grid=1001, every=10)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic",
                     constraint="ts1", lanes_per_instance=lanes, jacobian="monte_carlo_rev", jac_probes=10, jac_seed=1)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=False)


@pytest.mark.parametrize("fact,lanes,correct", [("isotropic", 0, True), ("isotropic", 16, False), ("dense", 0, True)])
def test_lml_value_with_solver_mle(oracle, pdq, fact, lanes, correct):
    """The LML value with solver_mle runs the cooperative LML on the posterior rescaled by the
    calibrated output scale (solvers.py:439-480); inputs where that scale moves the value.  Its
    gradient is refused."""
    params, y0, grid, every, data, stds = LC.lml_inputs_scale_sensitive(4, seed=17)
    cfg = pdq.Config(factorisation=fact, num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="mle",
                     lanes_per_instance=lanes, correct_asymptotic_underconfidence=correct)
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=False)
    with pytest.raises(pdq.PdqError, match="solver_mle"):
        pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds), grad=True)


def test_lml_tcoeff_index_is_checked(pdq):
    params, y0, grid, every, data, stds = LC.lml_inputs(1, seed=1, t_end=2.0, n_grid=11, every=5)
    for idx in (-1, 4):
        cfg = pdq.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", lml_tcoeff_index=idx)
        with pytest.raises(pdq.PdqError, match="lml_tcoeff_index"):
            pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds), grad=False)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", lanes_per_instance=1, lml_tcoeff_index=1)
    with pytest.raises(pdq.PdqError, match="lml_tcoeff_index = 0"):
        pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds), grad=False)


def test_lml_single_instance_value_only(oracle, pdq):
    params, y0, grid, every, data, stds = LC.lml_inputs(1, seed=9)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver")
    v_only, _ = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=False)
    v_grad, _ = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=True)
    # the gradient kernel finishes each step's backward conditional on its second wavefront
    # (lane_revert_right) and rounds differently from the value-only kernel: 1e-13 relative
    assert abs(v_only[0] - v_grad[0]) <= 1e-13 * abs(v_only[0])


def test_torch_autograd_function(oracle, pdq):
    from probdiffeq_amd import torch_ops

    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=4)
    prob = torch_ops.LMLProblem(vector_field="lotka_volterra", num_derivatives=3, ode_dim=2, grid=_dev(grid), every=every)
    P = _dev(params).requires_grad_(True)
    Y = _dev(y0).requires_grad_(True)
    value = torch_ops.lml(prob, _dev(data), _dev(stds), params=P, y0=Y)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (value * w).sum().backward()
    v_ref, g_ref = pdq.lml_value_and_grad(prob.config(4), _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))
    assert torch.equal(value.detach(), v_ref)
    full = torch.cat([P.grad, Y.grad], dim=1)
    assert torch.allclose(full, g_ref * w[:, None], rtol=0, atol=0)


def test_lane_lml_nu4_and_value_only(oracle, pdq):
    """The lane LML at nu = 4 (five Taylor coefficients), with and without the gradient."""
    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=12, t_end=10.0, n_grid=1001, every=10)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=4, ode_dim=2, strategy="fixedpoint", solver="solver", lanes_per_instance=1)
    v_grad, _ = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds)
    v_only, _ = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, grad=False)
    assert np.array_equal(v_only, v_grad)


def test_lane_lml_matches_the_cooperative_kernel_on_a_c5_batch(pdq):
    """At the C5 shape (2001 grid points, 101 observations), 64 instances: the lane LML and the
    cooperative kernel's LML agree to 1e-12 in value and 1e-9 in gradient (two FP64 orderings
    of the same algorithm; each is pinned to the oracle above at smaller sizes)."""
    params, y0, grid, every, data, stds = LC.lml_inputs(64, seed=21)
    base = dict(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver")
    args = (_dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))
    print("lml case: C5 shape, lane vs cooperative, B=64", flush=True)
    v1, g1 = pdq.lml_value_and_grad(pdq.Config(**base, lanes_per_instance=1), *args)
    v2, g2 = pdq.lml_value_and_grad(pdq.Config(**base, lanes_per_instance=16), *args)
    torch.cuda.synchronize()
    assert torch.isfinite(v1).all() and torch.isfinite(g1).all()
    assert ((v1 - v2).abs() / v2.abs()).max().item() < 1e-12
    assert ((g1 - g2).abs().amax(dim=1) / g2.abs().amax(dim=1)).max().item() < 1e-9


def test_lane_lml_refuses_other_shapes(pdq):
    params, y0, grid, every, data, stds = LC.lml_inputs(1, seed=2)
    cfg = pdq.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver", lanes_per_instance=1)
    with pytest.raises(pdq.PdqError, match="lane-per-instance LML"):
        pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))


def test_mirror_loss_lml_timeseries_observing_a_derivative(pdq):
    """probdiffeq.loss_lml_timeseries(tcoeff_index=1) through the torch op: value and autograd
    gradient equal the C ABI's with lml_tcoeff_index = 1, and differ from tcoeff_index = 0."""
    from probdiffeq_amd import probdiffeq as pd
    from probdiffeq_amd.torch_ops import LMLProblem

    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=12, t_end=5.0, n_grid=251, every=10)
    problem = LMLProblem(vector_field="lotka_volterra", num_derivatives=3, ode_dim=2, grid=_dev(grid), every=every,
                         factorisation="dense", solver="dynamic")
    p = _dev(params).requires_grad_(True)
    y = _dev(y0).requires_grad_(True)
    loss = pd.loss_lml_timeseries(tcoeff_index=1)
    v = loss(_dev(data), problem=problem, params=p, y0=y, std=_dev(stds))
    v.sum().backward()
    cfg = problem.config(params.shape[1])
    cfg.lml_tcoeff_index = 1
    vr, gr = pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))
    torch.cuda.synchronize()
    assert torch.equal(v.detach(), vr)
    assert torch.equal(torch.cat([p.grad, y.grad], dim=1), gr)
    v0 = pd.loss_lml_timeseries(tcoeff_index=0)(_dev(data), problem=problem, params=p.detach(), y0=y.detach(), std=_dev(stds))
    assert torch.all((v0 - v.detach()).abs() > 1.0)
    with pytest.raises(ValueError, match="tcoeff_index"):
        pd.loss_lml_timeseries(tcoeff_index=4)(_dev(data), problem=problem, params=p, y0=y, std=_dev(stds))


@pytest.mark.parametrize(
    "nu,solver,average,n_grid,every",
    [(3, "solver", True, 1001, 10), (3, "dynamic", False, 761, 20), (4, "solver", True, 1001, 10), (4, "dynamic", True, 761, 20),
     (3, "solver", True, 21, 5), (3, "dynamic", False, 51, 10), (3, "dynamic", True, 52, 3)],
)
def test_reverse_adjoint_against_finite_differences(oracle, pdq, nu, solver, average, n_grid, every):
    """lml_gradient = "reverse" (pdq_lane_adjoint.h: checkpoints every 50 grid steps, the forward
    factors recomputed per segment): the value against the oracle with the solve parity bound, the
    gradient against long double Richardson differences of the oracle to 1e-10.  761 grid points make
    the last checkpoint segment a partial one (760 = 15 x 50 + 10 steps); 21, 51 and 52 points give
    one partial segment, exactly one segment, and one segment plus a single step (the two-wavefront
    form's pipeline with one and two segments).  (On the 21-point grid solver_dynamic is left out:
    with the output scales held constant the LML is so curved there that the differences' step of
    1e-4 |x| misses the derivative, which they reach at 1e-6 |x|.)"""
    params, y0, grid, _, data, stds = LC.lml_inputs(3, seed=40 + nu, t_end=10.0, n_grid=n_grid, every=every)
    cfg = pdq.Config(factorisation="isotropic", num_derivatives=nu, ode_dim=2, strategy="fixedpoint", solver=solver, lml_gradient="reverse")
    _check(oracle, pdq, cfg, params, y0, grid, every, data, stds, average=average)


def test_reverse_adjoint_matches_forward_tangents_at_the_c5_shape(pdq):
    """The C5 workload (1024 instances, 2001 grid points, 101 observations, nu = 3): the reverse
    adjoint's value equals the forward-mode lane LML's to 1e-13 and its gradient the forward-mode
    tangents' to 1e-12, normwise relative per instance."""
    import bench

    params, y0, grid, data, stds = bench.c5_inputs(1024)
    every = 20
    base = dict(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver")
    args = (_dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))
    print("lml case: C5 shape, reverse vs forward, B=1024", flush=True)
    v_f, g_f = pdq.lml_value_and_grad(pdq.Config(**base), *args)
    v_r, g_r = pdq.lml_value_and_grad(pdq.Config(**base, lml_gradient="reverse"), *args)
    torch.cuda.synchronize()
    assert torch.isfinite(v_r).all() and torch.isfinite(g_r).all()
    assert ((v_r - v_f).abs() / v_f.abs()).max().item() < 1e-13
    err = ((g_r - g_f).abs().amax(dim=1) / g_f.abs().amax(dim=1)).max().item()
    print(f"reverse vs forward gradient: max normwise relative difference {err:.3e}", flush=True)
    assert err < 1e-12


def test_reverse_adjoint_value_only_and_instance_counts(pdq):
    """Without a gradient the reverse flag changes nothing (the forward LML runs); instance counts
    that do not fill a wavefront (1, 65) give each instance the value and gradient it has alone."""
    params, y0, grid, every, data, stds = LC.lml_inputs(65, seed=7, t_end=5.0, n_grid=251, every=10)
    base = dict(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic")
    args = (_dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))
    print("lml case: reverse adjoint, B=65 and B=1", flush=True)
    v0, _ = pdq.lml_value_and_grad(pdq.Config(**base), *args, grad=False)
    v1, _ = pdq.lml_value_and_grad(pdq.Config(**base, lml_gradient="reverse"), *args, grad=False)
    assert torch.equal(v0, v1)
    v65, g65 = pdq.lml_value_and_grad(pdq.Config(**base, lml_gradient="reverse"), *args)
    one = (_dev(params[64:]), _dev(y0[64:]), _dev(grid), every, _dev(data[64:]), _dev(stds[64:]))
    v1, g1 = pdq.lml_value_and_grad(pdq.Config(**base, lml_gradient="reverse"), *one)
    torch.cuda.synchronize()
    assert torch.equal(v65[64:], v1) and torch.equal(g65[64:], g1)


@pytest.mark.parametrize(
    "kw,match",
    [
        (dict(factorisation="dense"), "reverse adjoint"),
        (dict(num_derivatives=2), "reverse adjoint"),
        (dict(lml_tcoeff_index=1), "reverse adjoint"),
        (dict(lanes_per_instance=16), "one lane per instance"),
    ],
)
def test_reverse_adjoint_refuses_other_shapes(pdq, kw, match):
    params, y0, grid, every, data, stds = LC.lml_inputs(1, seed=2, t_end=5.0, n_grid=251, every=10)
    base = dict(factorisation="isotropic", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="solver", lml_gradient="reverse")
    base.update(kw)
    with pytest.raises(pdq.PdqError, match=match):
        pdq.lml_value_and_grad(pdq.Config(**base), _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds))


def test_lml_gain_solves_with_a_dropped_singular_value(oracle, pdq):
    """loss_lml_timeseries(solve_triu=...) on the device (dense, two observed rows per node): where
    R_Y's small singular value falls below lstsq_svd's cutoff the Jacobi-SVD gain matches the
    oracle's lstsq_svd and the back substitution the oracle's solve_triu (value, and gradient
    against Richardson differences to 1e-10); the two solves' likelihoods differ by more than 0.1 %."""
    params, y0, grid, every, data, stds = LC.lml_dropped_singular_value_inputs()
    vals = {}
    for solve in ("lstsq_svd", "solve_triu"):
        cfg = pdq.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic", lml_solve=solve)
        v, _ = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds)
        vals[solve] = v
    assert np.all(np.abs(vals["lstsq_svd"] - vals["solve_triu"]) > 1e-3 * np.abs(vals["solve_triu"]))


def test_mirror_loss_lml_timeseries_solve_choices(pdq):
    """probdiffeq.loss_lml_timeseries(solve_triu=linalg.solve_triu) reaches the device's back
    substitution (the value equals the C ABI's with lml_solve = "solve_triu"); another callable is
    refused."""
    from probdiffeq_amd import probdiffeq as pd
    from probdiffeq_amd.torch_ops import LMLProblem

    params, y0, grid, every, data, stds = LC.lml_dropped_singular_value_inputs()
    problem = LMLProblem(vector_field="lotka_volterra", num_derivatives=3, ode_dim=2, grid=_dev(grid), every=every,
                         factorisation="dense", solver="dynamic")
    for solve in (pd.linalg.lstsq_svd, pd.linalg.solve_triu):
        loss = pd.loss_lml_timeseries(solve_triu=solve)
        v = loss(_dev(data), problem=problem, params=_dev(params), y0=_dev(y0), std=_dev(stds))
        cfg = problem.config(params.shape[1])
        cfg.lml_solve = solve
        vr, _ = pdq.lml_value_and_grad(cfg, _dev(params), _dev(y0), _dev(grid), every, _dev(data), _dev(stds), grad=False)
        torch.cuda.synchronize()
        assert torch.equal(v.detach(), vr)
    with pytest.raises(NotImplementedError, match="solve_triu"):
        pd.loss_lml_timeseries(solve_triu=lambda a, b: b)


@pytest.mark.parametrize("solver", ["solver", "dynamic"])
def test_lml_with_an_exponential_prior(oracle, pdq, solver):
    """loss_lml_timeseries on a solve with a Matern prior (prior_matern, length scale 0.3; the
    dense exponential transition every step), through the C ABI and the torch op: value against
    the oracle, gradient against Richardson FD to 1e-10."""
    import bench
    from probdiffeq_amd.torch_ops import LMLProblem, lml

    params, y0, grid, every, data, stds = LC.lml_inputs(3, seed=6, t_end=5.0, n_grid=101, every=5, per_coordinate_std=True)
    drift = _dev(bench.matern_drift(0.3, 4, 2))
    cfg = pdq.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver=solver,
                     prior="exponential", prior_drift=drift)
    v, g = _check(oracle, pdq, cfg, params, y0, grid, every, data, stds)
    problem = LMLProblem(vector_field="lotka_volterra", num_derivatives=3, ode_dim=2, grid=_dev(grid), every=every,
                         factorisation="dense", solver=solver, prior="exponential", prior_drift=drift)
    P = _dev(params).requires_grad_(True)
    vt = lml(problem, _dev(data), _dev(stds), params=P, y0=_dev(y0))
    vt.sum().backward()
    assert torch.equal(vt.detach().cpu(), torch.as_tensor(v))
    assert torch.equal(P.grad.cpu(), torch.as_tensor(g[:, :4]))
