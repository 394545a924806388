"""partial_cmd_smooth prepared on the device (sfh_partial_cmd_smooth_batch):
mini_spacing, interpolate_mini, the callables, calculate_weights and midpoints
(src/StarFormationHistories.jl:77-148, 417-431, 765-776, 829-893) for many
isochrones in one call, then the deposition -- against the oracle's loop
restatement of the same preparation and its FP64 bin_cmd_smooth."""
import math

import numpy as np
import pytest

from helpers import max_norm_err, oracle_range

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _iso3(la, mh, n, seed, mmax):
    """A 3-filter isochrone: B, V and R = V - 0.4 (B - V)."""
    from sfh_amd import synth
    m, (B, V) = synth.isochrone(la, mh, n, seed, mmax_imf=mmax)
    return m, [B, V, V - 0.4 * (B - V)]


def _oracle_template(O, synth, m, mags, y_index, cidx, imf, edges, dmod, nv, compl=True):
    f = lambda v: float(synth.photerr(v))            # noqa: E731
    c = lambda v: float(synth.completeness(v))       # noqa: E731
    ref, pairs = O.partial_cmd_smooth(m, mags, [f] * len(mags), y_index, cidx, lambda v: float(imf(v)),
                                      [c] * len(mags) if compl else None, dmod=dmod, normalize_value=nv,
                                      mean_mass=imf.mean(), ex=oracle_range(O, edges[0]),
                                      ey=oracle_range(O, edges[1]), nthreads=8)
    return ref, pairs


@pytest.mark.parametrize("y_index,cidx", [(1, (0, 1)), (0, (0, 1)), (2, (0, 1))])
def test_device_prep_75x100(sfh, O, gpu_ctx, y_index, cidx):
    """Covariant (y in the colour, cov_mult +1 / -1) and asymmetric (y = R)
    templates of three isochrones with different distance moduli and masses,
    against the oracle's partial_cmd_smooth; the kernel-point and pair counts
    equal the oracle's preparation."""
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    isos = [_iso3(la, mh, 500, 21 + k, 2.0) for k, (la, mh) in enumerate([(9.2, -0.8), (8.3, -1.9), (10.0, 0.0)])]
    dmods = [24.0, 25.5, 23.2]
    nvs = [1e4, 3e3, 2e4]
    errs = [synth.PHOTERR] * 3
    comps = [synth.COMPLETENESS] * 3
    out, (npts, pairs) = sfh.partial_cmd_smooth_batch(isos, errs, y_index, cidx, imf, comps, dmod=dmods,
                                                      normalize_value=nvs, mean_mass=imf.mean(), edges=edges,
                                                      output="host", return_info=True)
    tot_pts, tot_pairs = 0, 0
    for t, (m, mags) in enumerate(isos):
        ref, pr = _oracle_template(O, synth, m, mags, y_index, cidx, imf, edges, dmods[t], nvs[t])
        got = out[:, t].reshape(ref.shape, order="F")
        assert max_norm_err(got, ref) <= TOL, t
        sp = O.partial_cmd_smooth_points(m, mags, [lambda v: float(synth.photerr(v))] * 3, y_index, cidx,
                                         lambda v: float(imf(v)), [lambda v: float(synth.completeness(v))] * 3,
                                         dmod=dmods[t], normalize_value=nvs[t], mean_mass=imf.mean(),
                                         ex=oracle_range(O, edges[0]), ey=oracle_range(O, edges[1]))
        tot_pts += sp.colors.size
        tot_pairs += pr
    assert npts == tot_pts and pairs == tot_pairs


def test_device_prep_config1_correlation(sfh, O, gpu_ctx):
    """The 700-template workload's combination (200 x 300, 2000 points, IMF
    0.08-120, correlation 0.6) for three isochrones: the device preparation
    against the host preparation (synth.template_points) deposited by the FP64
    oracle."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    specs = synth.specs_70x10()
    specs = [specs[3], specs[351], specs[-2]]
    out, (npts, pairs) = synth.templates_device(specs, 2000, edges, 120.0, correlation=0.6, output="host",
                                                return_info=True)
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    tot = 0
    for t, s in enumerate(specs):
        sp, cm = synth.template_points(s, 2000, edges, 120.0, correlation=0.6)
        ref, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey, cmult=cm,
                                   nthreads=8)
        tot += pr
        got = out[:, t].reshape(ref.shape, order="F")
        assert max_norm_err(got, ref) <= TOL, t
        big = ref > 1e-6 * ref.max()
        assert np.max(np.abs(got[big] - ref[big]) / ref[big]) <= 2e-5, t
    assert pairs == tot


def test_device_prep_outputs_and_tabulated_callables(sfh, gpu_ctx):
    """Device and host outputs agree bitwise; plain Python callables (tabulated
    on 16385 points) give the exact forms' templates to 1e-6; the template
    sum matches the weights' total inside the grid."""
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    isos = [synth.isochrone(la, -1.0, 400, 3, mmax_imf=2.0) for la in (9.0, 9.6)]
    kw = dict(dmod=24.0, normalize_value=1e4, mean_mass=imf.mean(), edges=edges)
    d = sfh.partial_cmd_smooth_batch(isos, [synth.PHOTERR] * 2, 1, (0, 1), imf, [synth.COMPLETENESS] * 2, **kw)
    h = sfh.partial_cmd_smooth_batch(isos, [synth.PHOTERR] * 2, 1, (0, 1), imf, [synth.COMPLETENESS] * 2,
                                     output="host", **kw)
    assert np.array_equal(d.cpu().numpy().T, h)
    # the packed input form gives the same templates bit for bit
    packed = sfh.partial_cmd_smooth_batch(sfh.IsochroneBatch.from_list(isos), [synth.PHOTERR] * 2, 1, (0, 1), imf,
                                          [synth.COMPLETENESS] * 2, output="host", **kw)
    assert np.array_equal(packed, h)
    f = lambda v: 0.01 + np.exp(0.6 * (np.asarray(v) - 28.0))           # noqa: E731
    c = lambda v: 1.0 / (1.0 + np.exp((np.asarray(v) - 28.0) / 0.7))    # noqa: E731
    tab = sfh.partial_cmd_smooth_batch(isos, [f] * 2, 1, (0, 1), lambda m: imf(m), [c] * 2, output="host", **kw)
    assert max_norm_err(tab, h) <= 1e-6


def test_device_prep_errors(sfh, gpu_ctx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    m, mags = synth.isochrone(9.0, -1.0, 300, 3, mmax_imf=2.0)
    args = ([synth.PHOTERR] * 2, 1, (0, 1), imf, [synth.COMPLETENESS] * 2)
    kw = dict(dmod=24.0, mean_mass=imf.mean(), edges=edges)
    with pytest.raises(sfh.ArgumentError):                       # unsorted masses
        sfh.partial_cmd_smooth_batch([(m[::-1], [v[::-1] for v in mags])], *args, **kw)
    with pytest.raises(sfh.ArgumentError):                       # a single point: no interpolation knots
        sfh.partial_cmd_smooth_batch([(m[:1], [v[:1] for v in mags])], *args, **kw)
    bad = [v.copy() for v in mags]
    bad[0][100] = np.nan
    with pytest.raises(sfh.InexactError):
        sfh.partial_cmd_smooth_batch([(m, bad)], *args, **kw)
    with pytest.raises(sfh.ArgumentError):
        sfh.partial_cmd_smooth_batch([(m, mags)], [synth.PHOTERR] * 2, 2, (0, 1), imf, **kw)
    # an empty isochrone next to a real one deposits nothing
    out = sfh.partial_cmd_smooth_batch([(m[:0], [v[:0] for v in mags]), (m, mags)], *args, output="host", **kw)
    assert not out[:, 0].any() and out[:, 1].any()
    assert math.isfinite(float(out.sum()))
