"""The reference's solver / sampler integration tests
(test/fitting/basic_linear_combinations.jl) with every objective evaluated
on the GPU.  Random templates come from numpy's seeded generator instead of
StableRNG, so the values differ but the checks are the reference's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def approx(a, b, rtol):
    """Julia's isapprox for arrays (the reference's `≈`): norm(a - b) <= rtol * max(norm(a), norm(b))."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def test_exact_single_template(sfh, gpu_ctx):
    # :16-37 -- one model, data = 3 * model, every optimiser recovers 3
    models = [np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1]], dtype=np.float64)]
    data = np.array([[0, 0, 0], [0, 0, 0], [3, 3, 3]], dtype=np.float64)
    tpl = sfh.HIPTemplates(models)
    d = data.ravel(order="F")
    assert sfh.fit_templates_lbfgsb(tpl, d, x0=[1.0])[1][0] == pytest.approx(3, rel=1e-7)
    assert sfh.fit_templates(tpl, d, x0=[1.0])["mle"].mu[0] == pytest.approx(3, rel=1e-7)
    assert sfh.fit_templates_fast(tpl, d, x0=[1.0])[0][0] == pytest.approx(3, rel=1e-7)


def test_noise_free_recovery(sfh, gpu_ctx):
    # :39-90 -- 10 random 100x100 templates, noise-free data, rtol 1e-7,
    # including coefficients that are exactly zero
    rng = np.random.default_rng(58392)
    T = 10
    x = rng.uniform(size=T)
    x0 = rng.uniform(size=T)
    models = rng.uniform(size=(100 * 100, T)).astype(np.float32)
    tpl = sfh.HIPTemplates(models)
    data = models.astype(np.float64) @ x
    assert approx(sfh.fit_templates_lbfgsb(tpl, data, x0=x0)[1], x, 1e-7)
    assert approx(sfh.fit_templates(tpl, data, x0=x0)["mle"].mu, x, 1e-7)
    assert approx(sfh.fit_templates_fast(tpl, data, x0=x0)[0], x, 1e-7)
    x2 = x.copy()
    x2[0] = x2[-1] = 0.0
    data2 = models.astype(np.float64) @ x2
    assert approx(sfh.fit_templates_lbfgsb(tpl, data2, x0=x0)[1], x2, 1e-7)
    assert approx(sfh.fit_templates(tpl, data2, x0=x0)["mle"].mu, x2, 1e-7)
    assert approx(sfh.fit_templates_fast(tpl, data2, x0=x0)[0], x2, 1e-7)


def test_poisson_recovery(sfh, gpu_ctx):
    # :91-118 -- Poisson data, rtol 1e-2 against truth; signatures agree
    rng = np.random.default_rng(94823)
    T = 10
    x = rng.uniform(size=T) * 100
    models = rng.uniform(size=(100 * 100, T)).astype(np.float32)
    tpl = sfh.HIPTemplates(models)
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    r1 = sfh.fit_templates_lbfgsb(tpl, data, x0=np.ones(T))[1]
    r2 = sfh.fit_templates(tpl, data, x0=np.ones(T))
    r3 = sfh.fit_templates_fast(tpl, data, x0=np.ones(T))[0]
    for r in (r1, r2["mle"].mu, r3):
        assert approx(r, x, 1e-2)
    assert approx(r2["mle"].mu, r3, 1e-5)
    assert r2["map"].sigma.shape == (T,) and np.all(r2["map"].sigma > 0)
    assert r2["map"].rand(5, rng=1).shape == (T, 5)


def test_hmc_model_and_batched(sfh, O, gpu_ctx):
    rng = np.random.default_rng(3)
    T, N = 6, 4000
    models = rng.uniform(size=(N, T)).astype(np.float32)
    x = rng.uniform(1, 5, T) * 10
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    hm = sfh.HMCModel(models, data)
    logx = np.log(x)
    ll, G = hm.logdensity_and_gradient(logx)
    assert ll == pytest.approx(hm.logdensity(logx), rel=1e-13)
    # finite differences of the log density
    for j in range(T):
        e = np.zeros(T)
        e[j] = 1e-6
        fd = (hm.logdensity(logx + e) - hm.logdensity(logx - e)) / 2e-6
        assert G[j] == pytest.approx(fd, rel=1e-5, abs=1e-5)
    LX = np.log(x)[:, None] + 0.1 * rng.standard_normal((T, 8))
    llb, Gb = hm.logdensity_and_gradient_batched(LX)
    for b in range(8):
        l1, g1 = hm.logdensity_and_gradient(LX[:, b])
        assert llb[b] == pytest.approx(l1, rel=1e-12)
        assert np.allclose(Gb[:, b], g1, rtol=1e-4, atol=1e-4 * np.abs(g1).max())


def test_mcmc_model_and_ensemble(sfh, gpu_ctx):
    rng = np.random.default_rng(5)
    T, N = 4, 3000
    models = rng.uniform(size=(N, T)).astype(np.float32)
    x = np.array([20.0, 40.0, 10.0, 30.0])
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    mm = sfh.MCMCModel(models, data)
    assert mm(-x) == -np.inf
    assert mm(x) == pytest.approx(sfh.loglikelihood(x, mm.models, mm.data), rel=1e-14)
    X0 = x[:, None] * (1 + 0.01 * rng.standard_normal((T, 16)))
    chain, lps, acc = sfh.mcmc_sample(mm.models, mm.data, X0, 60, rng=1)
    assert chain.shape == (60, T, 16) and lps.shape == (60, 16)
    assert 0.05 < acc < 0.95
    assert np.allclose(chain[30:].mean(axis=(0, 2)), x, rtol=0.2)
