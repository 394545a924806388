"""GPU parity of the binary-system templates (binary_hess(::RandomBinaryPairs),
src/StarFormationHistories.jl:627-706, and partial_cmd_smooth's binary branch
:895-910) against the oracle.  The pair histogram is accumulated in 64-bit
fixed point (absolute error per pair <= 2^-63 of a bound on the total), and
the completeness / bias callables are tabulated (16385 points, linear), so:
  pair histogram, callables constant : max-normalised error <= 1e-10
  pair histogram, logistic completeness and a bias : <= 1e-6
  templates (after bin_cmd_smooth)   : <= 1e-5, the template bound."""
import numpy as np
import pytest

from helpers import max_norm_err, oracle_range

pytestmark = pytest.mark.gpu


def _prep(sfh, y_index=1, cidx=(0, 1)):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    m, mags = synth.isochrone(9.2, -0.8, 500, 21, mmax_imf=2.0)
    sp = sfh.partial_cmd_smooth_points(m, mags, [synth.photerr] * 2, y_index, cidx, imf,
                                       [synth.completeness] * 2, dmod=24.0, normalize_value=1e4,
                                       mean_mass=imf.mean(), edges=edges)
    return synth, edges, imf, m, mags, sp


@pytest.mark.parametrize("callables", ["constant", "logistic+bias"])
def test_pair_histogram(sfh, O, gpu_ctx, callables):
    synth, edges, imf, m, mags, sp = _prep(sfh)
    mini = sp.new_mini[::4]
    iso = [v[::4] for v in sp.new_iso_mags]
    pdf = imf(mini)
    if callables == "constant":
        compl = [lambda v: np.ones_like(v)] * 2
        bias = [lambda v: np.zeros_like(v)] * 2
        ocompl, obias, tol = [lambda v: 1.0] * 2, [lambda v: 0.0] * 2, 1e-10
    else:
        compl = [synth.completeness] * 2
        bias = [lambda v: 0.01 * (v - 25.0), lambda v: -0.005 * (v - 25.0)]
        ocompl = [lambda v: float(synth.completeness(v))] * 2
        obias = [lambda v: 0.01 * (v - 25.0), lambda v: -0.005 * (v - 25.0)]
        tol = 1e-6
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    for y_is_a in (True, False):
        h = sfh.binary_pairs_hist(mini, pdf, iso[0], iso[1], y_is_a, 0.123, compl, bias, edges)
        ref = O.binary_pairs_hist(mini, pdf, iso[0], iso[1], y_is_a, 0.123, ocompl, obias, ex, ey)
        err = max_norm_err(h, ref)
        print(callables, y_is_a, f"pairs {mini.size - 1}: max-normalised error {err:.2e}")
        assert ref.sum() > 0 and err <= tol
        h2 = sfh.binary_pairs_hist(mini, pdf, iso[0], iso[1], y_is_a, 0.123, compl, bias, edges)
        assert np.array_equal(h, h2)          # fixed-point accumulation: bitwise repeatable


@pytest.mark.parametrize("y_index", [0, 1])
def test_binary_hess_and_partial_cmd_smooth(sfh, O, gpu_ctx, y_index):
    synth, edges, imf, m, mags, sp = _prep(sfh, y_index)
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    mini = sp.new_mini[::4]
    iso = [v[::4] for v in sp.new_iso_mags]
    model = sfh.RandomBinaryPairs(0.3)
    h = sfh.binary_hess(model, mini, iso, [synth.photerr] * 2, y_index, (0, 1), imf, [synth.completeness] * 2,
                        None, edges, normalize_value=1e4, mean_mass=imf.mean())
    ref = O.binary_hess(mini, iso, [lambda v: float(synth.photerr(v))] * 2, y_index, (0, 1),
                        lambda v: float(imf(v)), [lambda v: float(synth.completeness(v))] * 2, None,
                        ex=ex, ey=ey, normalize_value=1e4, mean_mass=imf.mean(), nthreads=8)
    err = max_norm_err(h.weights, ref)
    print(f"binary_hess y={y_index}: max-normalised error {err:.2e}")
    assert ref.sum() > 0 and err <= 1e-5
    full = sfh.partial_cmd_smooth(m, mags, [synth.photerr] * 2, y_index, (0, 1), imf, [synth.completeness] * 2,
                                  dmod=24.0, normalize_value=1e4, mean_mass=imf.mean(), edges=edges,
                                  binary_model=model)
    fref, _ = O.partial_cmd_smooth(m, mags, [lambda v: float(synth.photerr(v))] * 2, y_index, (0, 1),
                                   lambda v: float(imf(v)), [lambda v: float(synth.completeness(v))] * 2,
                                   dmod=24.0, normalize_value=1e4, mean_mass=imf.mean(), ex=ex, ey=ey,
                                   nthreads=8, binary_fraction=0.3)
    err = max_norm_err(full.weights, fref)
    print(f"partial_cmd_smooth + RandomBinaryPairs(0.3) y={y_index}: max-normalised error {err:.2e}")
    assert err <= 1e-5


def test_binary_model_edges(sfh, gpu_ctx):
    synth, edges, imf, m, mags, sp = _prep(sfh)
    single = sfh.partial_cmd_smooth(m, mags, [synth.photerr] * 2, 1, (0, 1), imf, dmod=24.0, edges=edges)
    none = sfh.partial_cmd_smooth(m, mags, [synth.photerr] * 2, 1, (0, 1), imf, dmod=24.0, edges=edges,
                                  binary_model=sfh.RandomBinaryPairs(0.0))
    assert np.array_equal(single.weights, none.weights)
    with pytest.raises(NotImplementedError):
        sfh.binary_hess(sfh.RandomBinaryPairs(0.5), sp.new_mini, sp.new_iso_mags + [sp.new_iso_mags[0]],
                        [synth.photerr] * 3, 2, (0, 1), imf, edges=edges)


def test_binary_hess_batch_matches_single_templates(sfh, gpu_ctx):
    """Many isochrones in one pair-histogram launch and one deposition launch
    give the per-template binary_hess results (tables shared over all the
    isochrones' magnitudes; FP32 templates), including an isochrone entirely
    off the grid (an empty template)."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    imf = synth.Kroupa2001(0.08, 120.0)
    isos = []
    for la, mh, dm in [(8.0, -1.0, 24.0), (9.0, -0.5, 25.0), (10.0, -2.0, 26.0), (9.5, -1.5, 45.0), (7.2, 0.0, 27.0)]:
        m, (B, V) = synth.isochrone(la, mh, 500, 3, mmax_imf=120.0)
        isos.append((m, [B + dm, V + dm]))
    errs = [synth.photerr, synth.photerr]
    comp = [synth.completeness, lambda v: synth.completeness(v, 0.4)]
    model = sfh.RandomBinaryPairs(0.3)
    T = sfh.binary_hess_batch(model, isos, errs, 1, (0, 1), imf, comp, None, edges,
                              normalize_value=1e3).cpu().numpy().astype(np.float64)
    H = sfh.binary_pairs_hist_batch(isos, imf, (0, 1), False,
                                    [1e3 / imf.mean() / 4 / sfh.trapz(m, imf(m)) for m, _ in isos], comp,
                                    [lambda v: 0 * v] * 2, edges)
    for t, (m, g) in enumerate(isos):
        ref = sfh.binary_hess(model, m, g, errs, 1, (0, 1), imf, comp, None, edges,
                              normalize_value=1e3).weights.ravel(order="F")
        href = sfh.binary_pairs_hist(m, imf(m), g[0], g[1], False, 1e3 / imf.mean() / 4 / sfh.trapz(m, imf(m)),
                                     comp, [lambda v: 0 * v] * 2, edges)
        if t == 3:
            assert np.all(T[t] == 0) and np.all(ref == 0) and np.all(H[t] == 0)
            continue
        # the shared tables span every isochrone (up to the one at dmod 45), so
        # their 16385 points are ~3x coarser than a single template's: the
        # linear interpolation of the steep completeness differs by ~1e-7
        assert np.max(np.abs(H[t] - href)) <= 1e-6 * np.max(np.abs(href))
        assert np.max(np.abs(T[t] - ref)) <= 1e-5 * np.max(np.abs(ref))
