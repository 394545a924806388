"""The plain method end to end on the 50-template workload: 10 log ages in
8.0-10.1 x 5 [M/H] in -2.0-0.0, 500-point isochrones, 75 x 100 Hess grid
(colour -0.5-2.5, mag 18-30), Kroupa IMF over 0.08-2 Msun, the exponential
photometric error and logistic completeness; the observed diagram a Poisson
draw of the composite with log-normal coefficients and ~1e5 stars; fitted with
fit_templates (src/fitting/solvers.jl:172-221).

Templates come from partial_cmd_smooth_batch (device preparation and
deposition) and are checked against the oracle's partial_cmd_smooth
(src/StarFormationHistories.jl:829-911); the fits -- the host BFGS around the
GPU fg! and the on-device BFGS of MultiFit -- are checked against SciPy BFGS
around the oracle's fg! on the same templates, and against the fit of the
oracle's own templates."""
import numpy as np
import pytest

from helpers import cpu_fit_templates, max_norm_err, oracle_range

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def config0(sfh, O, gpu_ctx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    specs = synth.specs_10x5()
    imf = synth.Kroupa2001(0.08, 2.0)
    out, (npts, pairs) = synth.templates_device(specs, 500, edges, 2.0, output="host", return_info=True)
    models = np.asfortranarray(out)                                  # 7500 x 50 FP32
    f = lambda v: float(synth.photerr(v))                            # noqa: E731
    c = lambda v: float(synth.completeness(v))                       # noqa: E731
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    ref = []
    for s in specs:
        m, mags = synth.isochrone(s.log_age, s.mh, 500, s.seed, mmax_imf=2.0)
        r, _ = O.partial_cmd_smooth(m, mags, [f, f], 1, (0, 1), lambda v: float(imf(v)), [c, c], dmod=24.0,
                                    normalize_value=1.0, mean_mass=imf.mean(), ex=ex, ey=ey, nthreads=8)
        ref.append(r.ravel(order="F"))
    ref = np.asfortranarray(np.stack(ref, axis=1))
    rng = np.random.default_rng(20261203)
    x_true = 10.0 ** (0.3 * rng.standard_normal(len(specs)))
    x_true *= 1e5 / (models.astype(np.float64) @ x_true).sum()
    data = synth.poisson_data(models.astype(np.float64) @ x_true, seed=5)
    return models, ref, data, x_true


def test_config0_templates_against_oracle(config0):
    models, ref, data, _ = config0
    assert models.shape == (7500, 50)
    for t in range(50):
        assert max_norm_err(models[:, t], ref[:, t]) <= 1e-5, t
    assert 0.95e5 <= data.sum() <= 1.05e5


def test_config0_fit_templates_against_cpu(sfh, O, config0):
    models, ref, data, x_true = config0
    gpu = sfh.fit_templates(models, data)
    cpu = cpu_fit_templates(O, models, data)
    for key in ("map", "mle"):
        mu_ref, f_ref = cpu[key]
        r = gpu[key]
        print(f"{key}: GPU fmin {r.result.fun:.12g} CPU {f_ref:.12g}")
        assert abs(r.result.fun - f_ref) <= 1e-9 * abs(f_ref)
        big = mu_ref > 1e-2 * mu_ref.max()
        assert np.allclose(r.mu[big], mu_ref[big], rtol=1e-3, atol=0), key
    # the fit of the oracle's own templates reaches the same optimum to the
    # templates' precision (FP32 deposition, max-normalised 1e-5)
    ref_fit = cpu_fit_templates(O, ref.astype(np.float32), data)
    f_oracle = ref_fit["mle"][1]
    print(f"MLE with the oracle's templates {f_oracle:.12g}")
    assert abs(gpu["mle"].result.fun - f_oracle) <= 1e-4 * abs(f_oracle) + 1e-3


def test_config0_device_fit(sfh, O, config0):
    """The same fit by the on-device BFGS loop (MultiFit, HIP-graph steps)."""
    from sfh_amd.multifit import MultiFit
    models, _, data, _ = config0
    res = MultiFit([models], [data]).fit_templates()[0]
    cpu = cpu_fit_templates(O, models, data)
    for key in ("map", "mle"):
        mu_ref, f_ref = cpu[key]
        r = res[key]
        print(key, r.result, "cpu", f_ref)
        assert r.result.status == "converged"
        assert abs(r.result.minimum - f_ref) <= 1e-9 * abs(f_ref)
        big = mu_ref > 1e-2 * mu_ref.max()
        assert np.allclose(r.mu[big], mu_ref[big], rtol=1e-3, atol=0), key
