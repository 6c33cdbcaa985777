"""The no-allocation solve (pdq_solve_adaptive_save_at_noalloc) against the
allocating entry point (pdq_solve_adaptive_save_at), through the C ABI.

Both run the same kernels on the same IWP matrices; only where the matrices
and the work-queue counter live differs (the library's fixed device workspace
instead of per-call and cached allocations), so every output must agree
bitwise.  Each case calls the no-allocation entry point twice in a row, the
second time on a smaller batch, so that the second call runs on the
workspace the first one left behind.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

THETA = np.array([0.5, 0.05, 0.5, 0.05])


@pytest.fixture(scope="module")
def pdq():
    from probdiffeq_amd import capi

    capi.require_device()
    return capi


def _dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


def _lv_batch(pdq, B, seed, nu, K=12):
    rng = np.random.default_rng(seed)
    params = THETA * np.exp(rng.uniform(np.log(0.8), np.log(1.2), size=(B, 4)))
    y0 = rng.uniform(10, 30, size=(B, 2))
    t_end = rng.uniform(4.0, 9.0, size=B)
    times = np.stack([np.linspace(0.0, te, K) for te in t_end])
    p, y = _dev(params), _dev(y0)
    return p, pdq.taylor_coefficients("lotka_volterra", p, y, 0.0, nu), _dev(times)


def _vdp_batch(pdq, B, seed, nu):
    rng = np.random.default_rng(seed)
    mu = _dev(np.exp(rng.uniform(np.log(1.0), np.log(30.0), size=(B, 1))))
    y0 = _dev(np.tile([2.0, 0.0], (B, 1)))
    return mu, pdq.taylor_coefficients("van_der_pol", mu, y0, 0.0, nu), _dev(np.tile(np.linspace(0.0, 3.0, 6), (B, 1)))


def _burgers_batch(pdq, B, seed, nu, Ng=100):
    rng = np.random.default_rng(seed)
    x = np.arange(Ng) / Ng
    y0 = np.stack([sum(rng.normal(0, 1 / k) * np.sin(2 * np.pi * k * x) for k in range(1, 9)) for _ in range(B)])
    p, y = _dev(np.tile([0.01, 1.0 / Ng], (B, 1))), _dev(y0)
    return p, pdq.taylor_coefficients("burgers", p, y, 0.0, nu), _dev(np.tile([0.0, 0.02, 0.05], (B, 1)))


CASES = {
    # dense filter, the C1 configuration: cooperative kernel
    "dense_filter": (dict(factorisation="dense", strategy="filter", solver="dynamic"), "lv", 4),
    # dense fixed-point smoother with the uncalibrated solver
    "dense_fixedpoint": (dict(factorisation="dense", num_derivatives=3, strategy="fixedpoint", solver="solver"), "lv", 3),
    # isotropic nu = 4, fixed-point, dynamic: the lane-per-instance kernel (C2 kind)
    "isotropic_lane": (dict(factorisation="isotropic", strategy="fixedpoint", solver="dynamic"), "lv", 4),
    # the same on the cooperative kernel with 8 lanes per instance
    "isotropic_coop": (dict(factorisation="isotropic", strategy="fixedinterval", solver="dynamic", lanes_per_instance=8), "lv", 4),
    # block-diagonal Burgers on 100 points: the block-diagonal kernel (one 128-lane group per instance)
    "blockdiag_burgers": (
        dict(factorisation="blockdiag", num_derivatives=2, ode_dim=100, vector_field="burgers", n_params=2, rtol=1e-4, atol=1e-6),
        "burgers",
        2,
    ),
    # dense ts1 on stiff van der Pol (C3 kind)
    "dense_ts1_vdp": (
        dict(factorisation="dense", num_derivatives=5, constraint="ts1", vector_field="van_der_pol", n_params=1, rtol=1e-5, atol=1e-7),
        "vdp",
        5,
    ),
}


@pytest.mark.parametrize("case", list(CASES))
def test_no_alloc_solve_equals_the_allocating_solve(pdq, case):
    kw, problem, nu = CASES[case]
    cfg = pdq.Config(**kw)
    for call, B in enumerate((12, 5)):
        seed = 100 * call + len(case)
        batch = {"lv": _lv_batch, "vdp": _vdp_batch, "burgers": _burgers_batch}[problem]
        p, tc, times = batch(pdq, B, seed, nu)
        ref = pdq.solve(cfg, p, tc, times)
        got = pdq.solve(cfg, p, tc, times, no_alloc=True)
        torch.cuda.synchronize()
        assert int(got["status"].abs().sum()) == 0, f"{case} call {call}: status {got['status'].tolist()}"
        assert set(got) == set(ref)
        for name in ref:
            assert torch.equal(got[name], ref[name]), f"{case} call {call} (B = {B}): {name} differs"


def test_no_alloc_solve_refuses_a_state_beyond_its_workspace(pdq):
    # dense nu = 7, d = 5: n = 40 > 32
    cfg = pdq.Config(factorisation="dense", num_derivatives=7, ode_dim=5, vector_field="linear", n_params=5)
    B, K = 2, 3
    p = _dev(np.full((B, 5), -0.5))
    tc = _dev(np.ones((B, 8, 5)))
    times = _dev(np.tile(np.linspace(0.0, 1.0, K), (B, 1)))
    with pytest.raises(pdq.PdqError, match="state dimension"):
        pdq.solve(cfg, p, tc, times, no_alloc=True)
