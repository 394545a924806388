"""Hierarchical fg! (src/fitting/hierarchical/amr.jl:82-171, mzr.jl:88-214)
with the likelihood on the GPU: the objective against the oracle's fg! at
the same coefficients, the gradient against central finite differences of
the objective (the reference pins its gradients with ForwardDiff the same
way, mzr_test.jl:66-74), the transformed log density
(generic_fitting.jl:100-199), and recovery of the truth by fit_sfh and the
fixed-weight fit on noise-free data."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(sfh, kind, N=3000, seed=3):
    from sfh_amd import synth
    H = sfh.hierarchical
    ula = np.arange(10.0, 8.95, -0.1)
    umh = np.arange(-2.0, 0.001, 0.25)
    logAge = np.repeat(ula, umh.size)
    MH = np.tile(umh, ula.size)
    models = synth.synthetic_templates(N, logAge.size, seed=seed)
    rng = np.random.default_rng(seed)
    mstars = rng.uniform(0.5, 1.5, ula.size) * 1e4
    if kind == "mzr":
        model = H.PowerLawMZR(0.6, -1.2, 4.0)
    elif kind == "linear_amr":
        model = H.LinearAMR(0.08, -1.8)
    else:
        model = H.LogarithmicAMR(0.0004, 0.0002)
    disp = H.GaussianDispersion(0.3)
    return H, models, logAge, MH, mstars, model, disp


def _fd(f, x, rel=1e-6):
    g = np.empty_like(x)
    for i in range(x.size):
        h = rel * max(abs(x[i]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("kind", ["mzr", "linear_amr", "log_amr"])
def test_fg_objective_and_gradient(sfh, O, gpu_ctx, kind):
    H, models, logAge, MH, mstars, model, disp = _grid(sfh, kind)
    coeffs = H.calculate_coeffs(model, disp, mstars, logAge, MH)
    data = np.random.default_rng(1).poisson(models.astype(np.float64) @ coeffs).astype(np.float64)
    tpl = sfh.HIPTemplates(models)
    hd = sfh.HIPData(data)
    v = np.concatenate([mstars, model.fittable_params(), disp.fittable_params()])
    G = np.empty_like(v)
    nll = H.fg_hierarchical(True, G, model, disp, v, tpl, hd, logAge, MH)
    rnl, _, _ = O.fg(models, data, coeffs, nthreads=8)
    assert abs(nll - rnl) <= 1e-12 * abs(rnl)
    assert H.fg_hierarchical(True, None, model, disp, v, tpl, hd, logAge, MH) == nll
    fd = _fd(lambda x: H.fg_hierarchical(True, None, model, disp, x, tpl, hd, logAge, MH), v)
    err = np.abs(G - fd) / np.maximum(np.abs(fd), 1e-6 * np.max(np.abs(fd)))
    print(kind, "max rel err vs finite differences", err.max())
    assert err.max() <= 1e-4
    # a fixed parameter gets no gradient entry
    fixed = model.update_params(model.fittable_params())
    fixed.free = (True, False)
    G2 = np.empty_like(v)
    H.fg_hierarchical(True, G2, fixed, disp, v, tpl, hd, logAge, MH)
    assert G2[mstars.size + 1] == 0.0 and np.allclose(G2[:mstars.size + 1], G[:mstars.size + 1], rtol=1e-12)


def test_logdensity_transformations(sfh, gpu_ctx):
    H, models, logAge, MH, mstars, model, disp = _grid(sfh, "mzr")
    coeffs = H.calculate_coeffs(model, disp, mstars, logAge, MH)
    data = np.random.default_rng(2).poisson(models.astype(np.float64) @ coeffs).astype(np.float64)
    tpl = sfh.HIPTemplates(models)
    hd = sfh.HIPData(data)
    xt = np.concatenate([np.log(mstars), [np.log(model.alpha), model.MH0, np.log(disp.sigma)]])
    v = np.concatenate([mstars, model.fittable_params(), disp.fittable_params()])
    nll = H.fg_hierarchical(True, None, model, disp, v, tpl, hd, logAge, MH)
    for jac in (False, True):
        prob = H.HierarchicalOptimizer(model, disp, tpl, hd, logAge, MH, jac)
        ll, g = prob.logdensity_and_gradient(xt)
        expect = -nll + (np.sum(np.log(mstars)) + np.log(model.alpha) + np.log(disp.sigma) if jac else 0.0)
        assert ll == pytest.approx(expect, rel=1e-12)       # mzr_test.jl:115-128
        fd = _fd(prob.logdensity, xt)
        assert np.max(np.abs(g - fd)) <= 1e-4 * np.max(np.abs(fd))
    # σ fixed: it leaves the variable vector (mzr_test.jl:131-140)
    dfix = H.GaussianDispersion(disp.sigma, (False,))
    prob = H.HierarchicalOptimizer(model, dfix, tpl, hd, logAge, MH, False)
    ll, g = prob.logdensity_and_gradient(xt[:-1])
    assert g.size == xt.size - 1 and ll == pytest.approx(-nll, rel=1e-12)


@pytest.mark.parametrize("kind", ["mzr", "linear_amr"])
def test_fit_sfh_noise_free_recovery(sfh, gpu_ctx, kind):
    H, models, logAge, MH, mstars, model, disp = _grid(sfh, kind, N=6000)
    coeffs = H.calculate_coeffs(model, disp, mstars, logAge, MH)
    data = models.astype(np.float64) @ coeffs
    start = model.update_params(np.array(model.fittable_params()) * np.array([1.3, 1.1]))
    res = sfh.hierarchical.fit_sfh(start, H.GaussianDispersion(0.4), models, data, logAge, MH,
                                   x0=mstars * 1.5)
    mle = res["mle"]
    print(kind, mle.mstars / mstars, mle.MH_params, model.fittable_params(), mle.disp_params)
    assert np.allclose(mle.mstars, mstars, rtol=1e-3)
    assert np.allclose(mle.MH_params, model.fittable_params(), rtol=1e-3, atol=1e-4)
    assert mle.disp_params[0] == pytest.approx(disp.sigma, rel=1e-3)
    assert np.all(res["map"].sigma > 0)


def test_fixed_amr_noise_free_recovery(sfh, gpu_ctx):
    H, models, logAge, MH, mstars, model, disp = _grid(sfh, "linear_amr", N=6000)
    rel = disp(MH, model(logAge))
    coeffs = H.calculate_coeffs(model, disp, mstars, logAge, MH)
    data = models.astype(np.float64) @ coeffs
    res = H.fit_templates_fixed_amr(models, data, logAge, MH, rel)
    assert np.allclose(res["mle"][0], mstars, rtol=1e-5)
