"""Parity of the drop-in boundary's finer semantics on the GPU:

* the two eps clamps of fg! (loglikelihood: eps(promote_type(composite, data)),
  src/fitting/fitting_base.jl:86; ∇loglikelihood!: eps(eltype(composite)), :277)
  through every fg! entry point (single pass, two pass, batched, grouped);
* the natural (vector-of-matrices) layout, composite!(::Matrix, coeffs,
  ::Vector{Matrix}) and ∇loglikelihood!(G, ::Matrix, ::Vector{Matrix}, ::Matrix)
  (:20-33, :227-256), with several templates and matrix-shaped composite/data;
* renormalize_x0's doctest (src/fitting/utilities.jl:89-102);
* HMCModel / MCMCModel log densities (src/fitting/hmc_sample.jl:11-37,
  src/fitting/mcmc_sample.jl:12-23) against the oracle's fg!.
"""
import numpy as np
import pytest

from helpers import ref_matrices

pytestmark = pytest.mark.gpu

E64 = float(np.finfo(np.float64).eps)
E32 = float(np.finfo(np.float32).eps)


def _tiny_bins_case(N, T, seed):
    """Templates whose composite is ~1e-10 (below eps(Float32), above
    eps(Float64)) on a band of bins that hold data, so the gradient depends on
    which clamp is used while the log-likelihood does not."""
    rng = np.random.default_rng(seed)
    models = rng.uniform(0.0, 1.0, (N, T)).astype(np.float32)
    band = slice(N // 3, N // 3 + 40)
    models[band, :] = rng.uniform(0.5e-10, 2e-10, (40, T)).astype(np.float32) / T
    x = rng.uniform(0.5, 2.0, T)
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    data[band] = rng.integers(1, 4, 40)
    return models, data, x


@pytest.mark.parametrize("N,T", [(3000, 12), (700, 800)])   # single-pass kernel / two-pass kernels
def test_fg_clamp_types(sfh, O, gpu_ctx, N, T):
    models, data, x = _tiny_bins_case(N, T, 11 + T)
    tpl = sfh.HIPTemplates(models)
    # composite::Vector{Float32} with Float64 data: LL eps(Float64), gradient eps(Float32)
    comp32 = np.empty(N, dtype=np.float32)
    G = np.empty(T)
    nl = sfh.fg_(True, G, x, tpl, data, comp32)
    rnl, rG, rcomp = O.fg(models, data, x, eps=E64, eps_grad=E32, nthreads=8)
    _, rG64, _ = O.fg(models, data, x, eps=E64, eps_grad=E64, nthreads=8)
    # the case discriminates: the two clamps give clearly different gradients
    assert np.max(np.abs(rG - rG64)) > 1e-3 * np.max(np.abs(rG64))
    assert nl == pytest.approx(rnl, rel=1e-12)
    assert np.max(np.abs(G - rG)) <= 1e-10 * np.max(np.abs(rG))
    # the clobbered composite is R = 1 - n / max(m, eps(Float32)) rounded to Float32
    assert np.allclose(comp32, rcomp.astype(np.float32), rtol=1e-6, atol=0)
    # composite::Vector{Float64}: both clamps eps(Float64)
    comp64 = np.empty(N)
    nl2 = sfh.fg_(True, G, x, tpl, data, comp64)
    assert nl2 == pytest.approx(rnl, rel=1e-12)
    assert np.max(np.abs(G - rG64)) <= 1e-10 * np.max(np.abs(rG64))
    # ∇loglikelihood! alone clamps with eps(eltype(composite)) (:277)
    c32 = np.asarray(models.astype(np.float64) @ x, dtype=np.float32)
    cc = c32.astype(np.float64)                 # before the call clobbers c32
    G3 = np.empty(T)
    sfh.grad_loglikelihood_(G3, c32, tpl, data)
    Gref = np.empty(T)
    O.grad_inplace(Gref, cc, models, data, eps=E32)
    assert np.max(np.abs(G3 - Gref)) <= 1e-10 * np.max(np.abs(Gref))


def test_driver_clamps(sfh):
    tpl32 = sfh.HIPTemplates(np.ones((10, 2), dtype=np.float32))
    tpl64 = sfh.HIPTemplates(np.ones((10, 2), dtype=np.float64))
    assert sfh.core.driver_clamps(tpl32, np.ones(10)) == (E64, E32)
    assert sfh.core.driver_clamps(tpl32, np.ones(10, dtype=np.int64)) == (E32, E32)
    assert sfh.core.driver_clamps(tpl64, np.ones(10, dtype=np.float32)) == (E64, E64)


@pytest.mark.parametrize("prec", ["f64", "fp32"])
def test_batched_clamp_types(sfh, O, gpu_ctx, prec):
    N, T, B = 4000, 24, 40
    models, data, x = _tiny_bins_case(N, T, 5)
    rng = np.random.default_rng(6)
    X = x[:, None] * 10.0 ** (0.1 * rng.standard_normal((T, B)))
    tpl = sfh.HIPTemplates(models)
    negll, G = sfh.fg_batched(X, tpl, data, precision=prec, eps=E64, eps_grad=E32)
    rnl, rG = O.fg_batched(models, data, X, eps=E64, eps_grad=E32, nthreads=8)
    _, rG64 = O.fg_batched(models, data, X, eps=E64, eps_grad=E64, nthreads=8)
    assert np.max(np.abs(rG - rG64)) > 1e-3 * np.max(np.abs(rG64))
    tol_ll, tol_g = (1e-12, 1e-6) if prec == "f64" else (1e-5, 1e-5)
    assert np.max(np.abs(negll - rnl) / np.abs(rnl)) <= tol_ll
    assert np.linalg.norm(G - rG) / np.linalg.norm(rG) <= tol_g


def test_multifit_clamp_types(sfh, O, gpu_ctx):
    fits = [_tiny_bins_case(3000, 10, 21), _tiny_bins_case(5000, 30, 22)]
    tpls = [sfh.HIPTemplates(m) for m, _, _ in fits]
    mf = sfh.MultiFit(tpls, [d for _, d, _ in fits])
    nl, G = mf.fg([x for _, _, x in fits], eps=E64, eps_grad=E32)
    for f, (m, d, x) in enumerate(fits):
        rnl, rG, _ = O.fg(m, d, x, eps=E64, eps_grad=E32, nthreads=8)
        assert nl[f] == pytest.approx(rnl, rel=1e-12)
        assert np.max(np.abs(G[f] - rG)) <= 1e-10 * np.max(np.abs(rG))


def test_natural_layout_multi_template(sfh, O, gpu_ctx):
    """composite!(::Matrix, coeffs, ::Vector{Matrix}) and the natural-layout
    ∇loglikelihood! (fitting_base.jl:20-33, :227-256) with five 30 x 40
    templates and matrix-shaped composite and data."""
    rng = np.random.default_rng(77)
    mats = [rng.uniform(0, 1, (30, 40)).astype(np.float32) for _ in range(5)]
    coeffs = rng.uniform(1, 5, 5)
    tpl = sfh.HIPTemplates(mats)
    assert tpl.shape == (1200, 5)
    comp = np.empty((30, 40))
    sfh.composite_(comp, coeffs, tpl)
    ref = sum(c * m.astype(np.float64) for c, m in zip(coeffs, mats))   # fill! + muladd passes (:27-31)
    assert np.max(np.abs(comp - ref)) <= 1e-14 * np.max(ref)
    data = rng.poisson(ref).astype(np.float64)
    assert sfh.loglikelihood(comp, data) == pytest.approx(O.loglikelihood(ref.ravel(order="F"), data.ravel(order="F")),
                                                         rel=1e-13)
    assert sfh.loglikelihood(coeffs, tpl, data) == pytest.approx(
        O.loglikelihood(ref.ravel(order="F"), data.ravel(order="F")), rel=1e-13)
    G = np.empty(5)
    sfh.grad_loglikelihood_(G, comp, tpl, data)
    # reference natural layout: composite <- 1 - n/max(m, eps(C)); G[k] = sum(-model_k .* composite)
    R = 1.0 - data / np.maximum(ref, E64)
    Gref = np.array([np.sum(-m.astype(np.float64) * R) for m in mats])
    assert np.max(np.abs(G - Gref)) <= 1e-11 * np.max(np.abs(Gref))
    assert comp.shape == (30, 40)
    assert np.max(np.abs(comp - R)) <= 1e-12 * np.max(np.abs(R))
    # the vector form on the same matrices, template by template (:171-182)
    g2 = sfh.grad_loglikelihood(tpl, ref, data)
    assert np.max(np.abs(g2 - Gref)) <= 1e-11 * np.max(np.abs(Gref))


def test_renormalize_x0_doctest(sfh, gpu_ctx):
    """fitting/utilities.jl:89-102 (rand(Float32, 5, 5) templates: the device
    stores FP32 templates)."""
    rng = np.random.default_rng(3)
    models = [rng.uniform(0, 1, (5, 5)).astype(np.float32) for _ in range(3)]
    true_x0 = np.array([1.0, 2.0, 3.0])
    data = sum(c * m.astype(np.float64) for c, m in zip(true_x0, models))     # noise-free
    x0 = true_x0 * 0.5
    x0_ren = sfh.renormalize_x0(data, models, x0)
    comp = sum(c * m.astype(np.float64) for c, m in zip(x0_ren, models))
    assert np.allclose(comp, data, rtol=np.sqrt(E64), atol=0)               # sum(x0_ren .* models) ≈ data
    assert np.allclose(x0_ren, true_x0, rtol=1e-14)
    # full_coeffs: the composite from full_coeffs, x0 scaled (:104-115)
    flat = sfh.stack_models(models)
    tpl = sfh.HIPTemplates(flat)
    y = sfh.renormalize_x0(data.ravel(order="F"), tpl, np.array([7.0]), true_x0 * 0.25)
    assert y == pytest.approx([28.0], rel=1e-14)
    # degenerate composite: x0 returned unchanged (:111)
    z = sfh.renormalize_x0(data.ravel(order="F"), tpl, np.zeros(3))
    assert np.array_equal(z, np.zeros(3))


def test_hmc_mcmc_model_values(sfh, O, gpu_ctx):
    """HMCModel: logdensity = logL(exp θ) + Σθ and its gradient G .* x .+ 1
    (hmc_sample.jl:11-37); MCMCModel: logL(θ), -Inf for any θ < 0
    (mcmc_sample.jl:12-23); against the oracle's fg!."""
    rng = np.random.default_rng(8)
    N, T = 5000, 20
    models = rng.uniform(0, 1, (N, T)).astype(np.float32)
    x = rng.uniform(1, 3, T)
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    m = sfh.HMCModel(models, data)
    theta = np.log(x * rng.uniform(0.8, 1.2, T))
    nl, G, _ = O.fg(models, data, np.exp(theta), eps=E64, eps_grad=E32, nthreads=8)
    assert m.logdensity(theta) == pytest.approx(-nl + theta.sum(), rel=1e-13)
    ld, g = m.logdensity_and_gradient(theta)
    assert ld == pytest.approx(-nl + theta.sum(), rel=1e-13)
    gref = -G * np.exp(theta) + 1.0
    assert np.max(np.abs(g - gref)) <= 1e-10 * np.max(np.abs(gref))
    # the reference's 3x3 known answer through HMCModel: θ = log.(coeffs)
    mats, d3, c3 = ref_matrices()
    h3 = sfh.HMCModel(mats, d3)
    assert h3.logdensity(np.log(c3)) == pytest.approx(-1.4180233783775342 + np.log(c3).sum(), rel=1e-14)
    mc = sfh.MCMCModel(models, data)
    rl = O.loglikelihood(models.astype(np.float64) @ x, data)
    assert mc(x) == pytest.approx(rl, rel=1e-13)
    xb = x.copy()
    xb[3] = -1e-3
    assert mc(xb) == -np.inf
