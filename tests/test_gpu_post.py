"""GPU: loss_lml_terminal_values (estimators_and_losses.py:20-50) through the C ABI and the mirror API,
against the oracle (pinned to scipy's Gaussian log-density in tests/test_oracle_offgrid.py)."""

import numpy as np
import pytest

import test_gpu_parity as T

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda()


@pytest.mark.parametrize("nu,idx", [(4, 0), (3, 2)])
def test_lml_terminal_values_match_the_oracle(oracle, nu, idx):
    from probdiffeq_amd import capi

    capi.require_device()
    rng = np.random.default_rng(nu + idx)
    B, N, d = 16, nu + 1, 2
    mean = rng.normal(size=(B, N, d))
    chol = np.tril(rng.normal(size=(B, N, N))) + 2.0 * np.eye(N)
    u = rng.normal(size=(B, d))
    std = rng.uniform(0.1, 1.0, size=B)
    cfg = capi.Config(factorisation="isotropic", num_derivatives=nu, ode_dim=d)
    print(f"terminal lml case: nu={nu} idx={idx} B={B}", flush=True)
    got = capi.lml_terminal_values(cfg, _dev(mean), _dev(chol), _dev(u), _dev(std), idx).cpu().numpy()
    ocfg = oracle.Config(fact="isotropic", nu=nu, d=d, strategy="filter")
    for b in range(B):
        vals = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                vals.append(float(oracle.lml_terminal(ocfg, T.THETA, mean[b].ravel(), chol[b].ravel(), u[b], std[b], idx)))
        T.check_accuracy(got[b], *vals, f"terminal lml nu={nu} idx={idx} instance {b}")


@pytest.mark.parametrize("nu,idx,d", [(3, 0, 2), (4, 2, 2), (2, 1, 3)])
def test_dense_lml_terminal_values_match_the_oracle(oracle, nu, idx, d):
    """The dense form: derivative idx's d coordinates observed with noise diag(std), one QR of the
    stacked factor, the d-dimensional Gaussian log-density (the oracle pinned to scipy)."""
    from probdiffeq_amd import capi

    capi.require_device()
    rng = np.random.default_rng(10 * nu + idx)
    B, n = 12, (nu + 1) * d
    mean = rng.normal(size=(B, n))
    chol = np.tril(rng.normal(size=(B, n, n))) + 2.0 * np.eye(n)
    u = rng.normal(size=(B, d))
    std = rng.uniform(0.1, 1.0, size=(B, d))
    cfg = capi.Config(factorisation="dense", num_derivatives=nu, ode_dim=d)
    print(f"dense terminal lml case: nu={nu} idx={idx} d={d} B={B}", flush=True)
    got = capi.lml_terminal_values(cfg, _dev(mean), _dev(chol), _dev(u), _dev(std), idx).cpu().numpy()
    ocfg = oracle.Config(fact="dense", nu=nu, d=d, strategy="filter")
    for b in range(B):
        vals = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                vals.append(float(oracle.lml_terminal(ocfg, T.THETA, mean[b], chol[b].ravel(), u[b], std[b], idx)))
        T.check_accuracy(got[b], *vals, f"dense terminal lml nu={nu} idx={idx} instance {b}")


def test_api_lml_terminal_values_of_a_terminal_value_solve(oracle):
    import test_gpu_api as A

    from probdiffeq_amd import ivpsolve
    from probdiffeq_amd import probdiffeq as pdx

    y0 = torch.tensor([[20.0, 20.0], [15.0, 25.0]], dtype=torch.float64, device="cuda")
    init, slv, err = A._setup(pdx, "isotropic", "filter", "dynamic", y0)
    sol = ivpsolve.solve_adaptive_terminal_values(solver=slv, error=err)(init, t0=0.0, t1=5.0, atol=1e-8, rtol=1e-6)
    data = torch.tensor([[10.0, 30.0], [12.0, 28.0]], dtype=torch.float64, device="cuda")
    std = torch.tensor([0.5, 0.5], dtype=torch.float64, device="cuda")
    print("terminal lml api case: B=2", flush=True)
    value = pdx.loss_lml_terminal_values()(data, marginals=sol.u, std=std)
    assert value.shape == (2,)
    ocfg = oracle.Config(fact="isotropic", nu=4, d=2, strategy="filter")
    for b in range(2):
        m = sol.u.mean_flat[b].cpu().numpy().ravel()
        L = sol.u.cholesky_flat[b].cpu().numpy().ravel()
        vals = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                vals.append(float(oracle.lml_terminal(ocfg, T.THETA, m, L, data[b].cpu().numpy(), 0.5, 0)))
        T.check_accuracy(value[b].item(), *vals, f"terminal lml api instance {b}")


def test_api_dense_lml_terminal_values(oracle):
    """The mirror on a dense terminal-value solve: std is shaped as marginals.std[0] (d,) per instance."""
    import test_gpu_api as A

    from probdiffeq_amd import ivpsolve
    from probdiffeq_amd import probdiffeq as pdx

    y0 = torch.tensor([[20.0, 20.0], [15.0, 25.0]], dtype=torch.float64, device="cuda")
    init, slv, err = A._setup(pdx, "dense", "filter", "dynamic", y0)
    sol = ivpsolve.solve_adaptive_terminal_values(solver=slv, error=err)(init, t0=0.0, t1=5.0, atol=1e-8, rtol=1e-6)
    data = torch.tensor([[10.0, 30.0], [12.0, 28.0]], dtype=torch.float64, device="cuda")
    std = torch.tensor([[0.5, 0.3], [0.4, 0.6]], dtype=torch.float64, device="cuda")
    print("dense terminal lml api case: B=2", flush=True)
    with pytest.raises(ValueError, match="standard deviation container"):
        pdx.loss_lml_terminal_values()(data, marginals=sol.u, std=std[:, 0])
    value = pdx.loss_lml_terminal_values()(data, marginals=sol.u, std=std)
    ocfg = oracle.Config(fact="dense", nu=4, d=2, strategy="filter")
    for b in range(2):
        m = sol.u.mean_flat[b].cpu().numpy().ravel()
        L = sol.u.cholesky_flat[b].cpu().numpy().ravel()
        vals = []
        for build in ("strict", "fma", "extended"):
            with oracle.rounding(build):
                vals.append(float(oracle.lml_terminal(ocfg, T.THETA, m, L, data[b].cpu().numpy(), std[b].cpu().numpy(), 0)))
        T.check_accuracy(value[b].item(), *vals, f"dense terminal lml api instance {b}")


@pytest.mark.parametrize("nu", [3, 4])
def test_posterior_samples_match_the_oracle(oracle, nu):
    """MarkovSequence.sample (estimators_and_losses.py:233-273) through the C ABI, with the oracle's
    draws, against the oracle in its three builds."""
    import sample_cases as SC

    from probdiffeq_amd import capi

    capi.require_device()
    cfg, sol, base, refs = SC.make_case(oracle, nu)
    K = base.shape[1]
    print(f"sample case: nu={nu} S={base.shape[0]} K={K}", flush=True)
    out = {k: _dev(v[None]) for k, v in sol.items()}
    samples = capi.sample(cfg, out, K, _dev(base[None]))
    torch.cuda.synchronize()
    SC.check(samples[0].cpu().numpy(), refs, f"gpu sample nu={nu}")


def _matern_drift(length_scale, N, d):
    from math import comb

    z = np.sqrt(2 * (N - 0.5)) / length_scale
    return np.concatenate([-comb(N, i) * z ** (N - i) * np.eye(d) for i in range(N)], axis=1)


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
def test_generic_sampler_matches_the_oracle(oracle, fact, nu, strategy, kw):
    """Sampling for the shapes without a lane sampler (dense with n draws per step; other isotropic
    shapes) through the C ABI, against the oracle in its three builds."""
    import sample_cases as SC

    from probdiffeq_amd import capi

    capi.require_device()
    cfg, sol, base, refs = SC.make_case(oracle, nu, fact=fact, strategy=strategy, **kw)
    if cfg.prior_drift is not None:
        cfg.prior_drift = _dev(cfg.prior_drift)
    K = base.shape[1]
    print(f"generic sample case: {fact} nu={nu} {strategy} S={base.shape[0]} K={K}", flush=True)
    out = {k: _dev(v[None]) for k, v in sol.items()}
    samples = capi.sample(cfg, out, K, _dev(base[None]))
    torch.cuda.synchronize()
    SC.check(samples[0].cpu().numpy(), refs, f"gpu generic sample {fact}/{strategy} nu={nu}")


def test_api_posterior_samples():
    """The mirror's MarkovSequence.sample: shapes, a seed's reproducibility, and the draws path equal
    to the C ABI's; a filter's solution has no posterior to sample."""
    import test_gpu_api as A

    from probdiffeq_amd import ivpsolve
    from probdiffeq_amd import probdiffeq as pdx

    y0 = torch.tensor([[20.0, 20.0], [15.0, 25.0], [25.0, 12.0]], dtype=torch.float64, device="cuda")
    init, slv, err = A._setup(pdx, "isotropic", "fixedpoint", "dynamic", y0)
    save_at = torch.linspace(0.0, 10.0, 30, dtype=torch.float64, device="cuda")
    sol = ivpsolve.solve_adaptive_save_at(solver=slv, error=err)(init, save_at=save_at, atol=1e-8, rtol=1e-6)
    post = sol.solution_full.posterior
    print("sample api case: B=3 K=30", flush=True)
    s1 = post.sample(7, shape=(4,))
    s2 = post.sample(7, shape=(4,))
    assert s1.shape == (3, 4, 30, 5, 2)
    assert torch.equal(s1, s2)
    assert post.sample(3).shape == (3, 30, 5, 2)
    base = torch.randn(3, 2, 30, 5, dtype=torch.float64, device="cuda")
    via_api = post.sample_from_draws(base)
    from probdiffeq_amd import capi

    out = {
        "term_mean": post.marginal.mean_flat.reshape(3, 10).contiguous(),
        "term_chol": post.marginal.cholesky_flat.reshape(3, 25).contiguous(),
        "cond_A": post.conditional.A.reshape(3, 29, 25).contiguous(),
        "cond_mean": post.conditional.noise.mean_flat.reshape(3, 29, 10).contiguous(),
        "cond_chol": post.conditional.noise.cholesky_flat.reshape(3, 29, 25).contiguous(),
        "cond_tl": post.conditional.to_latent.reshape(3, 29, 5).contiguous(),
        "cond_to": post.conditional.to_observed.reshape(3, 29, 5).contiguous(),
    }
    direct = capi.sample(sol.config, out, 30, base)
    assert torch.equal(via_api.reshape(3, 2, 30, 10), direct)
    # the terminal sample row is the terminal marginal's mean plus its factor times the draw
    lz = torch.matmul(post.marginal.cholesky_flat[:, None].contiguous(), base[:, :, -1, :, None].contiguous())[..., 0]
    exp = post.marginal.mean_flat[:, None] + lz[..., None]
    assert torch.allclose(via_api[:, :, -1], exp, rtol=1e-14, atol=1e-12)


def test_api_posterior_samples_dense(oracle):
    """The mirror's MarkovSequence.sample on a dense fixed-point solve: shapes (K, N d), a seed's
    reproducibility, and each sample equal to the oracle's with the same draws."""
    import test_gpu_api as A

    from probdiffeq_amd import ivpsolve
    from probdiffeq_amd import probdiffeq as pdx

    y0 = torch.tensor([[20.0, 20.0], [15.0, 25.0]], dtype=torch.float64, device="cuda")
    init, slv, err = A._setup(pdx, "dense", "fixedpoint", "dynamic", y0)
    save_at = torch.linspace(0.0, 6.0, 12, dtype=torch.float64, device="cuda")
    sol = ivpsolve.solve_adaptive_save_at(solver=slv, error=err)(init, save_at=save_at, atol=1e-8, rtol=1e-6)
    post = sol.solution_full.posterior
    print("dense sample api case: B=2 K=12", flush=True)
    s1 = post.sample(5, shape=(3,))
    assert s1.shape == (2, 3, 12, 10)
    assert torch.equal(s1, post.sample(5, shape=(3,)))
    base = torch.randn(2, 2, 12, 10, dtype=torch.float64, device="cuda")
    got = post.sample_from_draws(base).cpu().numpy()
    ocfg = T._oracle_cfg(oracle, sol.config)
    c = post.conditional
    for b in range(2):
        arrs = {
            "term_mean": post.marginal.mean_flat[b].cpu().numpy(),
            "term_chol": post.marginal.cholesky_flat[b].cpu().numpy(),
            "cond_A": c.A[b].cpu().numpy(),
            "cond_mean": c.noise.mean_flat[b].cpu().numpy(),
            "cond_chol": c.noise.cholesky_flat[b].cpu().numpy(),
            "cond_tl": c.to_latent[b].cpu().numpy(),
            "cond_to": c.to_observed[b].cpu().numpy(),
        }
        for s in range(2):
            z = base[b, s].cpu().numpy()
            per = []
            for build in ("strict", "fma", "extended"):
                with oracle.rounding(build):
                    per.append(np.asarray(oracle.sample(ocfg, T.THETA, arrs, z), dtype=np.float64))
            T.check_accuracy(got[b, s], *per, f"dense api sample {b} {s}")
