"""GPU parity of the single-vector likelihood path (composite!, loglikelihood,
∇loglikelihood(!), fg!) against the oracle and the reference's known answers."""
import numpy as np
import pytest

from helpers import golden, ref_matrices, stack

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def test_known_answers(sfh, O, gpu_ctx):
    models, data, coeffs = ref_matrices()
    tpl = sfh.HIPTemplates(stack(models))
    d = data.ravel(order="F")
    # composite! is exact (fitting_core_test.jl:20,28)
    comp = np.zeros(9)
    sfh.composite_(comp, coeffs, tpl)
    assert np.array_equal(comp, O.composite(stack(models), coeffs))
    # loglikelihood(composite, data) (:39)
    C = np.array([[1, 1, 1], [2, 2, 2], [3, 3, 3]], dtype=np.float64)
    assert sfh.loglikelihood(C, data) == pytest.approx(-0.5672093513510137, rel=1e-14)
    Cz = np.array([[1.5] * 3, [3] * 3, [3] * 3])
    dz = np.array([[0] * 3, [2] * 3, [2] * 3], dtype=np.float64)
    assert sfh.loglikelihood(Cz, dz) == pytest.approx(-5.6344187027020260, rel=1e-14)
    # fg! (:176-192)
    G = np.empty(3)
    nl = sfh.fg_(True, G, coeffs, tpl, d)
    assert nl == pytest.approx(1.4180233783775342, rel=1e-14)
    assert np.allclose(-G, [-1, -1, -1], rtol=1e-14)
    assert sfh.fg_(True, None, coeffs, tpl, d) == pytest.approx(1.4180233783775342, rel=1e-14)
    # ∇loglikelihood! with zero data (:149-159)
    G3 = np.empty(3)
    comp3 = O.composite(stack(models), coeffs)
    sfh.grad_loglikelihood_(G3, comp3, tpl, dz.ravel(order="F"))
    assert np.allclose(G3, [-3, -1, -1], rtol=1e-14)
    assert np.allclose(comp3, 1 - dz.ravel(order="F") / O.composite(stack(models), coeffs), rtol=1e-15)
    # ∇loglikelihood(coeffs, models, data) (:113-121)
    assert np.allclose(sfh.grad_loglikelihood(coeffs, tpl, d), [-1, -1, -1], rtol=1e-14)


def test_mpmath_golden(sfh, gpu_ctx):
    g = golden("likelihood.json")
    N, T = g["nbins"], g["ntemplates"]
    m = np.array(g["models_colmajor"], dtype=np.float32).reshape((N, T), order="F")
    tpl = sfh.HIPTemplates(m)
    comp = np.empty(N)
    sfh.composite_(comp, g["coeffs"], tpl)
    assert np.allclose(comp, g["composite"], rtol=1e-14)
    assert sfh.loglikelihood(g["coeffs"], tpl, g["data"]) == pytest.approx(g["loglikelihood"], rel=1e-13)
    G = sfh.grad_loglikelihood(g["coeffs"], tpl, g["data"])
    assert np.allclose(G, g["grad"], rtol=1e-12)


# 700 and below run fg! in one pass over the templates; 900 takes the two-pass kernels
@pytest.mark.parametrize("N,T", [(60000, 700), (7500, 50), (1, 1), (257, 3), (100003, 13), (20000, 900), (3000, 704)])
def test_fg_against_oracle(sfh, O, gpu_ctx, N, T):
    from sfh_amd import synth
    models = synth.synthetic_templates(N, T, seed=N + T)
    rng = np.random.default_rng(T)
    x = rng.uniform(0.1, 2.0, T) * (1e6 / T)
    data = synth.poisson_data(models.astype(np.float64) @ x, seed=7)
    tpl = sfh.HIPTemplates(models)
    hd = sfh.HIPData(data)
    G = np.empty(T)
    comp = np.empty(N)
    nl = sfh.fg_(True, G, x, tpl, hd, comp)
    rnl, rG, rcomp = O.fg(models, data, x, nthreads=8)
    assert nl == pytest.approx(rnl, rel=1e-12)
    assert np.max(np.abs(G - rG)) <= 1e-10 * np.max(np.abs(rG)) + 1e-300
    # the clobbered composite R = 1 - n/c: absolute error ~ ulp(n/c)
    assert np.max(np.abs(comp - rcomp)) <= 1e-13 * max(1.0, np.max(np.abs(1.0 - rcomp)))


@pytest.mark.parametrize("T", [40, 700, 900])
def test_fg_output_combinations(sfh, O, gpu_ctx, T):
    """fg! with F only, G only and both agree; the composite output holds m
    without a gradient and R = 1 - n/max(m, eps) with one (solvers.jl:26-37)."""
    from sfh_amd import synth
    N = 9000
    models = synth.synthetic_templates(N, T, seed=T)
    x = np.random.default_rng(T).uniform(0.1, 2.0, T) * (1e5 / T)
    data = synth.poisson_data(models.astype(np.float64) @ x, seed=3)
    tpl = sfh.HIPTemplates(models)
    hd = sfh.HIPData(data)
    G = np.empty(T)
    comp_R = np.empty(N)
    nl = sfh.fg_(True, G, x, tpl, hd, comp_R)
    G2 = np.empty(T)
    # the G-only branch returns G, the value of its last expression `G .*= -1` (solvers.jl:31-33)
    assert sfh.fg_(None, G2, x, tpl, hd) is G2
    assert np.array_equal(G, G2)
    comp_m = np.empty(N)
    assert sfh.fg_(True, None, x, tpl, hd, comp_m) == nl
    m = O.composite(models, x)
    assert np.max(np.abs(comp_m - m)) <= 1e-13 * np.max(m)
    rnl, rG, rcomp = O.fg(models, data, x, nthreads=8)
    assert nl == pytest.approx(rnl, rel=1e-12)
    assert np.max(np.abs(G - rG)) <= 1e-10 * np.max(np.abs(rG))
    assert np.max(np.abs(comp_R - rcomp)) <= 1e-13 * max(1.0, np.max(np.abs(1.0 - rcomp)))
    assert sfh.loglikelihood(x, tpl, hd) == -nl


def test_device_path_matches_host_path(sfh, O, gpu_ctx):
    import torch
    from sfh_amd import synth
    N, T = 30000, 64
    models = synth.synthetic_templates(N, T, seed=1)
    x = np.random.default_rng(1).uniform(1e3, 1e4, T)
    data = synth.poisson_data(models.astype(np.float64) @ x, seed=2)
    tpl = sfh.HIPTemplates(models)
    hd = sfh.HIPData(data)
    dx = torch.from_numpy(x).cuda()
    dnl = torch.empty(1, dtype=torch.float64, device="cuda")
    dG = torch.empty(T, dtype=torch.float64, device="cuda")
    sfh.fg_device(tpl, hd, dx, dnl, dG)
    gpu_ctx.synchronize()
    G = np.empty(T)
    nl = sfh.fg_(True, G, x, tpl, hd)
    assert dnl.item() == nl and np.array_equal(dG.cpu().numpy(), G)
    # repeated calls are bitwise identical (fixed-order reductions)
    G2 = np.empty(T)
    assert sfh.fg_(True, G2, x, tpl, hd) == nl and np.array_equal(G, G2)


def test_zero_sum_is_minus_inf(sfh, gpu_ctx):
    assert sfh.loglikelihood(np.array([2.0, 3.0]), np.array([2.0, 3.0])) == -np.inf


def test_float32_eps_semantics(sfh, O, gpu_ctx):
    # eps(promote_type(Float32, Int)) = eps(Float32) clamps a zero model bin
    c = np.array([0.0, 1.0], dtype=np.float32)
    d = np.array([2, 1], dtype=np.int64)
    e32 = float(np.finfo(np.float32).eps)
    assert sfh.loglikelihood(c, d) == pytest.approx(O.loglikelihood(c, d, eps=e32), rel=1e-14)


def test_argument_errors(sfh, gpu_ctx):
    tpl = sfh.HIPTemplates(np.ones((10, 3), dtype=np.float32))
    with pytest.raises(sfh.ArgumentError):
        sfh.fg_(True, np.empty(3), np.ones(3), tpl, np.ones(11))
    with pytest.raises(sfh.ArgumentError):
        sfh.fg_(True, np.empty(3), np.ones(4), tpl, np.ones(10))
    with pytest.raises(sfh.ArgumentError):
        sfh.loglikelihood(np.ones(3), np.ones(4))


def test_templates_from_device_pointer(sfh, O, gpu_ctx):
    import torch
    from sfh_amd import synth
    N, T = 5000, 9
    models = synth.synthetic_templates(N, T, seed=3)
    dev = torch.from_numpy(np.ascontiguousarray(models.T)).cuda()     # T x N row-major == N x T col-major
    tpl = sfh.HIPTemplates(device_ptr=dev.data_ptr(), shape=(N, T))
    x = np.linspace(1, 2, T)
    comp = np.empty(N)
    sfh.composite_(comp, x, tpl)
    assert np.allclose(comp, O.composite(models, x), rtol=1e-14)
