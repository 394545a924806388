"""Many independent fits of uneven size on the GPU (sfh_multifit_*): the
grouped fg! against the FP64 oracle fit by fit, and the device optimiser's
fit_templates / fit_templates_fast against the same fits done on the CPU
(SciPy BFGS around the oracle) and against the reference's recovery tests
(test/fitting/basic_linear_combinations.jl)."""
import numpy as np
import pytest

from helpers import cpu_fit_templates

pytestmark = pytest.mark.gpu


def approx(a, b, rtol):
    """Julia's isapprox for arrays: norm(a - b) <= rtol * max(norm(a), norm(b))."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def _fit_problem(N, T, seed, total=None):
    from sfh_amd import synth
    models = synth.synthetic_templates(N, T, seed=seed)
    rng = np.random.default_rng(seed + 7)
    x = rng.lognormal(0.0, 1.0, T)
    if total is not None:
        x *= total / (models.astype(np.float64) @ x).sum()
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    return models, data, x


SHAPES = [(100, 1), (31, 7), (5000, 50), (12345, 300), (20000, 700), (257, 1024), (64, 33)]


def test_grouped_fg_against_oracle(sfh, O, gpu_ctx):
    from sfh_amd.multifit import MultiFit
    probs = [_fit_problem(N, T, seed=10 + i, total=1e5) for i, (N, T) in enumerate(SHAPES)]
    mf = MultiFit([p[0] for p in probs], [p[1] for p in probs])
    rng = np.random.default_rng(5)
    X = [p[2] * np.exp(0.3 * rng.normal(size=p[2].size)) for p in probs]
    # the group's largest fit (1024 templates) takes 4-wave workgroups
    shapes = [(1 << 18, 4, 2), (1 << 16, 4, 3), (4096, 4, 1)]
    results = []
    for shp in shapes:
        mf.set_shape(*shp)
        nl, G = mf.fg(X)
        results.append((nl, G))
        for f, (models, data, _) in enumerate(probs):
            rnl, rG, _ = O.fg(models, data, X[f], nthreads=8)
            assert abs(nl[f] - rnl) <= 1e-12 * abs(rnl), (shp, f, nl[f], rnl)
            assert np.linalg.norm(G[f] - rG) <= 1e-10 * np.linalg.norm(rG), (shp, f)
    # every work decomposition gives the same numbers to rounding
    for nl, G in results[1:]:
        assert np.allclose(nl, results[0][0], rtol=1e-13, atol=0)
    print("autotuned shape:", mf.autotune(reps=2))
    nl, _ = mf.fg(X)
    assert np.allclose(nl, results[0][0], rtol=1e-13, atol=0)
    # 8-wave workgroups take groups whose fits have at most 704 templates,
    # 16-wave ones at most 512
    for waves in (8, 16):
        with pytest.raises(sfh.ArgumentError):
            mf.set_shape(1 << 20, waves, 1)
    for limit, waves_shapes in ((704, [(1 << 18, 8, 2), (4096, 8, 1)]), (512, [(1 << 20, 16, 1), (1 << 16, 16, 2)])):
        small = [i for i, (_, T) in enumerate(SHAPES) if T <= limit]
        ms = MultiFit([probs[i][0] for i in small], [probs[i][1] for i in small])
        for shp in waves_shapes:
            ms.set_shape(*shp)
            nl, G = ms.fg([X[i] for i in small])
            for j, i in enumerate(small):
                rnl, rG, _ = O.fg(probs[i][0], probs[i][1], X[i], nthreads=8)
                assert abs(nl[j] - rnl) <= 1e-12 * abs(rnl), (shp, i)
                assert np.linalg.norm(G[j] - rG) <= 1e-10 * np.linalg.norm(rG), (shp, i)
        ms.close()


def test_grouped_fg_repeatable_and_without_gradient(sfh, gpu_ctx):
    from sfh_amd.multifit import MultiFit
    probs = [_fit_problem(N, T, seed=40 + i) for i, (N, T) in enumerate([(3000, 20), (777, 5)])]
    mf = MultiFit([p[0] for p in probs], [p[1] for p in probs])
    X = [p[2] for p in probs]
    a, Ga = mf.fg(X)
    b, Gb = mf.fg(X)
    c, Gc = mf.fg(X, want_G=False)
    assert np.array_equal(a, b) and all(np.array_equal(u, v) for u, v in zip(Ga, Gb))
    assert np.array_equal(a, c) and Gc is None


def test_fit_templates_against_cpu(sfh, O, gpu_ctx):
    from sfh_amd.multifit import MultiFit
    shapes = [(500, 3), (1000, 8), (2000, 15), (3000, 20)]
    probs = [_fit_problem(N, T, seed=70 + i, total=2e4) for i, (N, T) in enumerate(shapes)]
    mf = MultiFit([p[0] for p in probs], [p[1] for p in probs])
    res = mf.fit_templates()
    for f, (models, data, _) in enumerate(probs):
        ref = cpu_fit_templates(O, models, data)
        for key in ("map", "mle"):
            r = res[f][key]
            mu_ref, fmin_ref = ref[key]
            print(f, key, r.result, "cpu fmin", fmin_ref)
            assert r.result.status == "converged"
            big = mu_ref > 1e-3 * mu_ref.max()
            assert np.allclose(r.mu[big], mu_ref[big], rtol=1e-5, atol=0), (f, key)
            assert abs(r.result.minimum - fmin_ref) <= 1e-9 * abs(fmin_ref) + 1e-9, (f, key)
            H = r.invH
            assert H.shape == (shapes[f][1],) * 2
            assert np.max(np.abs(H - H.T)) <= 1e-10 * np.max(np.abs(H))
            assert np.all(np.diag(H) > 0) and np.all(r.sigma > 0)


def test_fit_templates_fast_and_lbfgs_agree(sfh, O, gpu_ctx):
    """fit_templates_fast (θ² transform) with BFGS and L-BFGS against the CPU
    fit, and log-space L-BFGS at the reference's suggested g_abstol = 1e-3
    (solvers.jl:232): the same MLE as BFGS.  (In log space a coefficient whose
    MLE is 0 drifts to θ -> -inf with curvature ~ e^θ, which one L-BFGS scale
    cannot follow at g_abstol = 1e-8; dense BFGS can.)"""
    from sfh_amd.multifit import MultiFit
    probs = [_fit_problem(N, T, seed=90 + i, total=5e4) for i, (N, T) in enumerate([(4000, 12), (1500, 30)])]
    mf = MultiFit([p[0] for p in probs], [p[1] for p in probs])
    log_b = mf.fit_templates(method="bfgs")
    log_l = mf.fit_templates(method="lbfgs", g_abstol=1e-3)
    fast_b = mf.fit_templates_fast(method="bfgs")
    fast_l = mf.fit_templates_fast(method="lbfgs")
    for f, (models, data, _) in enumerate(probs):
        ref_mu, ref_f = cpu_fit_templates(O, models, data, transform="sqrt")["mle"]
        mu = log_b[f]["mle"].mu
        big = mu > 1e-3 * mu.max()
        print(f, "bfgs", log_b[f]["mle"].result, "lbfgs", log_l[f]["mle"].result)
        print(f, "fast bfgs", fast_b[f][1], "fast lbfgs", fast_l[f][1], "cpu", ref_f)
        # BFGS: converged, the CPU optimum to rounding
        assert fast_b[f][1].status == "converged" and log_b[f]["mle"].result.status == "converged"
        assert np.allclose(fast_b[f][0][big], ref_mu[big], rtol=1e-5, atol=0)
        assert abs(fast_b[f][1].minimum - ref_f) <= 1e-9 * abs(ref_f)
        assert np.allclose(fast_b[f][0][big], mu[big], rtol=1e-5, atol=0)
        # L-BFGS: slower on the ill-conditioned fit (it may stop at the
        # 1000-iteration cap with |g| just above 1e-8) but at the same optimum
        coeffs, info = fast_l[f]
        assert info.status in ("converged", "max_iterations")
        assert abs(info.minimum - ref_f) <= 1e-9 * abs(ref_f)
        assert np.allclose(coeffs[big], ref_mu[big], rtol=1e-4, atol=0)
        # log-space L-BFGS stalls on the zero-MLE coefficients of fit 1 (see
        # above): only its objective is checked
        assert log_l[f]["mle"].result.status in ("converged", "max_iterations")
        assert log_l[f]["mle"].invH is None
        assert abs(log_l[f]["mle"].result.minimum - ref_f) <= 1e-5 * abs(ref_f)


def test_reference_recovery_through_the_group(sfh, gpu_ctx):
    """basic_linear_combinations.jl:16-118 as one group of fits: exact single
    template (x = 3), noise-free recovery (rtol 1e-7, with exact zeros) and
    Poisson recovery (rtol 1e-2)."""
    from sfh_amd.multifit import MultiFit
    m1 = np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1]], dtype=np.float32).ravel(order="F")[:, None]
    d1 = np.array([[0, 0, 0], [0, 0, 0], [3, 3, 3]], dtype=np.float64).ravel(order="F")
    rng = np.random.default_rng(58392)
    x = rng.uniform(size=10)
    x0 = rng.uniform(size=10)
    m2 = rng.uniform(size=(100 * 100, 10)).astype(np.float32)
    d2 = m2.astype(np.float64) @ x
    x3 = x.copy()
    x3[0] = x3[-1] = 0.0
    d3 = m2.astype(np.float64) @ x3
    rng = np.random.default_rng(94823)
    xp = rng.uniform(size=10) * 100
    m4 = rng.uniform(size=(100 * 100, 10)).astype(np.float32)
    d4 = rng.poisson(m4.astype(np.float64) @ xp).astype(np.float64)
    mf = MultiFit([m1, m2, m2, m4], [d1, d2, d3, d4])
    x0s = [np.ones(1), x0, x0, np.ones(10)]
    res = mf.fit_templates(x0=x0s)
    fast = mf.fit_templates_fast(x0=x0s)
    assert res[0]["mle"].mu[0] == pytest.approx(3, rel=1e-7)
    assert fast[0][0][0] == pytest.approx(3, rel=1e-7)
    assert approx(res[1]["mle"].mu, x, 1e-7) and approx(fast[1][0], x, 1e-7)
    assert approx(res[2]["mle"].mu, x3, 1e-7) and approx(fast[2][0], x3, 1e-7)
    assert approx(res[3]["mle"].mu, xp, 1e-2) and approx(fast[3][0], xp, 1e-2)
    assert approx(res[3]["mle"].mu, fast[3][0], 1e-5)


def test_graph_length_does_not_change_results(sfh, gpu_ctx):
    from sfh_amd.multifit import MultiFit
    probs = [_fit_problem(N, T, seed=120 + i, total=1e4) for i, (N, T) in enumerate([(800, 6), (2500, 25)])]
    mf = MultiFit([p[0] for p in probs], [p[1] for p in probs])
    a = mf.fit_templates(steps_per_graph=1)
    b = mf.fit_templates(steps_per_graph=32)
    for f in range(2):
        for key in ("map", "mle"):
            assert np.array_equal(a[f][key].mu, b[f][key].mu)
            assert a[f][key].result.iterations == b[f][key].result.iterations


def test_multifit_argument_errors(sfh, gpu_ctx):
    from sfh_amd.multifit import MultiFit
    with pytest.raises(sfh.ArgumentError):
        MultiFit([np.ones((10, 1025), dtype=np.float32)], [np.ones(10)])
    with pytest.raises(sfh.ArgumentError):
        MultiFit([np.ones((10, 3), dtype=np.float32)], [np.ones(11)])
    mf = MultiFit([np.ones((10, 3), dtype=np.float32)], [np.ones(10)])
    with pytest.raises(sfh.ArgumentError):
        mf.fg([np.ones(4)])
    with pytest.raises(sfh.ArgumentError):
        mf.set_shape(1 << 16, 5, 2)
    with pytest.raises(sfh.ArgumentError):
        mf.fit_templates(method="lbfgs", memory=0)
