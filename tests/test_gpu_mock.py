"""The GPU mock-observation path (sfh_generate_stars, sfh_model_cmd,
sfh_bin_cmd) against the oracle restatement (oracle/mock_ref.py) drawing
the same counter-based random numbers -- the same systems, the same kept
entries, masses and magnitudes to rounding -- plus the reference's own
statistical checks (test/cmd_simulation_test.jl) through the GPU API and a
Monte-Carlo cross-check of partial_cmd_smooth: a population sampled, observed
and binned on the GPU against the smoothed template of the same isochrone."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mock_ref as M   # noqa: E402

pytestmark = pytest.mark.gpu

M_INI = np.round(np.arange(1, 9) / 10.0, 10)
F606W = np.array([12.73830341586291, 10.017833567325411, 9.041398536300997, 8.331313172834694, 7.3892413746605765,
                  6.227971374447669, 4.93799980882831, -3.026])
F814W = np.array([11.137576129331675, 8.951605387408511, 8.064408199510426, 7.424658866406447, 6.6025729402403,
                  5.584714238650148, 4.448999828801471, -4.107])
NAMES = ["F606W", "F814W"]


def _oracle_imf(imf):
    return M.PowerLaw(imf.breaks, imf.slopes, imf.coefs)


def _synth_iso(log_age=9.0, mh=-1.0):
    from sfh_amd import synth
    m, (B, V) = synth.isochrone(log_age, mh, 2000, 7, mmax_imf=120.0)
    return m, np.stack([B, V], axis=1)


def _bkind(model, sfh):
    return 0 if isinstance(model, sfh.NoBinaries) else (1 if isinstance(model, sfh.RandomBinaryPairs) else 2)


def _compare(sfh, mini, mags, names, mode, limit, imf, model, dmod, mag_lim, mag_lim_name, seed, stream,
             limit_name=None):
    from sfh_amd import mock
    if mode == 0:
        cat = mock.generate_stars_mass(mini, mags, names, limit, imf, dist_mod=dmod, seed=seed, stream=stream,
                                       mag_lim=mag_lim, mag_lim_name=mag_lim_name, binary_model=model, catalog=True)
    else:
        cat = mock.generate_stars_mag(mini, mags, names, limit, limit_name, imf, dist_mod=dmod, seed=seed,
                                      stream=stream, mag_lim=mag_lim, mag_lim_name=mag_lim_name, binary_model=model,
                                      catalog=True)
    gm, gg, gi = cat.arrays()
    g = mock.ingest_mags(mini, mags) + dmod
    m, g = mock.sort_ingested(mini, g)
    kws = dict(seed=seed, stream=stream)
    if mode == 0:
        cut, _ = mock.mass_limits(m, g, names, mag_lim, mag_lim_name)
        kws["mmin_cut"] = cut
        lim = limit
    else:
        lim = 10 ** (4 * (0 - (limit + dmod)) / 10)
        kws["limit_filter"] = names.index(limit_name)
        if np.isfinite(mag_lim):
            kws.update(mag_lim=mag_lim, mag_lim_filter=names.index(mag_lim_name))
    qm = model.qmin if hasattr(model, "qmin") else 0.0
    qx = model.qmax if hasattr(model, "qmax") else 1.0
    om, og, oi, drawn = M.generate(m, g, mode, lim, _oracle_imf(imf), _bkind(model, sfh),
                                   getattr(model, "fraction", 0.0), qm, qx, **kws)
    print(f"mode {mode} {type(model).__name__}: {drawn} systems drawn, {gi.size} kept")
    assert cat.sampled == drawn
    assert np.array_equal(gi, oi)
    # pow-based IMF inversion: near the upper mass limit the two terms of
    # m = (lo^a + y a)^(1/a) cancel (22.4 - 0.002 for Salpeter at 100 Msun),
    # magnifying the 1-ulp difference of device and host pow up to ~1e4
    assert np.allclose(gm, om, rtol=1e-10, atol=0)
    fin = np.isfinite(og)
    assert np.array_equal(np.isfinite(gg), fin)
    assert np.max(np.abs(gg[fin] - og[fin]), initial=0.0) < 1e-11
    cat.close()
    return gm, gg


@pytest.mark.parametrize("model", ["none", "pairs", "ratio"])
def test_generate_mass_matches_oracle(sfh, gpu_ctx, model):
    from sfh_amd import mock
    bm = {"none": sfh.NoBinaries(), "pairs": sfh.RandomBinaryPairs(0.4), "ratio": mock.BinaryMassRatio(0.4)}[model]
    imf = mock.Salpeter1955(0.1, 100.0)
    _compare(sfh, M_INI, [F606W, F814W], NAMES, 0, 1e5, imf, bm, 20.0, math.inf, "F814W", 11, 1)
    # a real isochrone, a Kroupa IMF and a magnitude limit (mass_limits cut)
    m, g = _synth_iso()
    _compare(sfh, m, g, ["B", "V"], 0, 3e5, mock.Kroupa2001(0.08, 120.0), bm, 24.0, 28.5, "V", 5, 2)


@pytest.mark.parametrize("model", ["none", "pairs", "ratio"])
def test_generate_mag_matches_oracle(sfh, gpu_ctx, model):
    from sfh_amd import mock
    bm = {"none": sfh.NoBinaries(), "pairs": sfh.RandomBinaryPairs(0.4), "ratio": mock.BinaryMassRatio(0.4)}[model]
    imf = mock.Salpeter1955(0.1, 100.0)
    _compare(sfh, M_INI, [F606W, F814W], NAMES, 1, -8.0, imf, bm, 20.0, math.inf, "F814W", 12, 3, "F814W")
    _compare(sfh, M_INI, [F606W, F814W], NAMES, 1, -8.0, imf, bm, 20.0, 31.0, "F606W", 12, 4, "F814W")
    # many batches: a population far larger than the first batch of 2^20 systems
    m, g = _synth_iso()
    _compare(sfh, m, g, ["B", "V"], 1, -15.0, mock.Kroupa2001(0.08, 120.0), bm, 24.0, math.inf, "V", 6, 5, "V")


def test_reference_statistics_through_gpu(sfh, gpu_ctx):
    """test/cmd_simulation_test.jl:126-214 on the GPU path."""
    from scipy import integrate
    from sfh_amd import mock
    imf = mock.Salpeter1955(0.1, 100.0)
    f = lambda m: m * m ** -2.35
    mass_frac = integrate.quad(f, 0.1, 0.8)[0] / integrate.quad(f, 0.1, 100.0)[0]
    mags = [F606W, F814W]
    for bm, factor in ((sfh.NoBinaries(), 1.0), (sfh.RandomBinaryPairs(0.4), 1.4), (mock.BinaryMassRatio(0.4), 1.0)):
        ms, g = sfh.generate_stars_mass(M_INI, mags, NAMES, 1e5, imf, dist_mod=20.0, seed=21, binary_model=bm)
        assert ms.shape[1] == (1 if isinstance(bm, sfh.NoBinaries) else 2) and g.shape == (ms.shape[0], 2)
        assert ms.sum() == pytest.approx(1e5 * mass_frac * factor, rel=5e-2)
        ms, g = sfh.generate_stars_mag(M_INI, mags, NAMES, -8.0, "F814W", imf, dist_mod=20.0, seed=22,
                                       binary_model=bm)
        absmag = sfh.flux2mag(np.sum(sfh.mag2flux(g[:, 1]))) - 20.0
        assert -0.05 <= absmag + 8.0 <= 0.0
    with pytest.raises(sfh.ArgumentError):
        sfh.generate_stars_mass(M_INI, mags, NAMES, 1e5, imf, mag_lim=25.0, mag_lim_name="V")
    with pytest.raises(sfh.ArgumentError):
        sfh.generate_stars_mag(M_INI, mags, NAMES, -8.0, "V", imf)
    with pytest.raises(sfh.ArgumentError):
        sfh.generate_stars_mag(M_INI, mags, NAMES, -8.0, "F814W", imf, mag_lim=25.0, mag_lim_name="V")
    # a luminosity no system can reach (IMF above every isochrone mass) is an error, not an endless loop
    with pytest.raises(sfh.DomainError):
        sfh.generate_stars_mag(M_INI, mags, NAMES, -8.0, "F814W", mock.Salpeter1955(1.0, 100.0),
                               binary_model=sfh.NoBinaries())
    # composite: two spoofed isochrones, half the mass each (:219-238)
    cm = [M_INI, M_INI + 0.01]
    cg = [mags, [F606W, F814W - 0.02]]
    rm, rg = sfh.generate_stars_mass_composite(cm, cg, NAMES, 1e5, [0.5, 0.5], imf, dist_mod=20.0, seed=23,
                                               binary_model=sfh.NoBinaries())
    assert len(rm) == len(rg) == 2 and all(a.shape[0] == b.shape[0] for a, b in zip(rm, rg))
    assert sum(a.sum() for a in rm) == pytest.approx(1e5 * mass_frac, rel=5e-2)
    rm, rg = sfh.generate_stars_mag_composite(cm, cg, NAMES, -8.0, "F814W", [0.5, 0.5], imf, dist_mod=20.0,
                                              seed=24, binary_model=sfh.NoBinaries())
    tot = sum(np.sum(sfh.mag2flux(b[:, 1] - 20.0)) for b in rg)
    assert sfh.mag2flux(-8.0) <= tot <= sfh.mag2flux(-8.0 - 0.05)


def test_model_cmd_matches_oracle(sfh, gpu_ctx):
    from sfh_amd import mock, synth
    rng = np.random.default_rng(3)
    mags = np.stack([rng.uniform(19, 30, 300_000), rng.uniform(18, 29, 300_000)], axis=1)
    errs = [synth.photerr, lambda m: 0.5 * synth.photerr(m)]
    comps = [synth.completeness, lambda m: synth.completeness(m, 0.4)]
    bias = [lambda m: 0.01 * (m - 25.0), lambda m: np.zeros_like(m)]
    out, idx = sfh.model_cmd(mags, errs, comps, bias, seed=9, stream=2, ret_idxs=True)
    lo, hi = mags.min(axis=0), mags.max(axis=0)

    def tabs(fs):
        res = []
        for f, fn in enumerate(fs):
            grid = np.linspace(lo[f], hi[f], 16385)
            res.append((grid[0], grid[1] - grid[0], np.asarray(fn(grid), dtype=np.float64)))
        return res
    ro, ri = M.model_cmd(mags, tabs(errs), tabs(comps), tabs(bias), seed=9, stream=2)
    print(f"model_cmd: kept {idx.size} of {mags.shape[0]}")
    assert np.array_equal(idx, ri)
    assert np.max(np.abs(out - ro)) < 1e-12
    # the tabulated callables against the callables themselves
    e = out - mags[idx]
    assert abs(np.mean(e[:, 1])) < 5 * np.mean(0.5 * synth.photerr(mags[idx, 1])) / math.sqrt(idx.size) + 1e-4
    with pytest.raises(sfh.ArgumentError):
        sfh.model_cmd(mags, errs[:1], comps)
    # the tables' magnitude range comes from the device and skips non-finite entries
    cat = mock.Catalog.from_mags(np.array([[20.0, np.inf], [22.0, 19.0], [21.0, 18.5]]))
    lo, hi = cat.mag_range()
    assert lo.tolist() == [20.0, 18.5] and hi.tolist() == [22.0, 19.0]
    cat.close()


def test_bin_cmd_matches_oracle(sfh, gpu_ctx):
    from sfh_amd import mock, synth
    rng = np.random.default_rng(4)
    x, y, w = rng.uniform(-1, 3, 400_000), rng.uniform(17, 31, 400_000), rng.uniform(0, 3, 400_000)
    ex, ey = synth.edges_200x300()
    h = sfh.bin_cmd(x, y, edges=(ex, ey))
    ref = M.bin_cmd(x, y, ex.first, ex.step, ex.length - 1, ey.first, ey.step, ey.length - 1)
    assert np.array_equal(h.weights, ref) and h.closed == "left"
    hw = sfh.bin_cmd(x, y, weights=w, edges=(ex, ey))
    refw = M.bin_cmd(x, y, ex.first, ex.step, ex.length - 1, ey.first, ey.step, ey.length - 1, weights=w)
    assert np.max(np.abs(hw.weights - refw)) <= 1e-12 * np.max(refw)
    # from a device catalogue (colour = filter 0 - filter 1, y = filter 1)
    cat = mock.Catalog.from_mags(np.stack([x + y, y], axis=1))
    hc = mock.bin_cmd_catalog(cat, (0, 1), 1, (ex, ey))
    xc = (x + y) - y
    assert np.array_equal(hc.weights, M.bin_cmd(xc, y, ex.first, ex.step, ex.length - 1, ey.first, ey.step,
                                                ey.length - 1))
    cat.close()
    with pytest.raises(sfh.ArgumentError):
        sfh.bin_cmd(x, y[:-1], edges=(ex, ey))


def test_monte_carlo_population_matches_template(sfh, gpu_ctx):
    """A 3e5 Msun population of one isochrone, sampled, observed (errors and
    completeness of the template) and binned on the GPU, against
    partial_cmd_smooth of that isochrone for the same birth mass."""
    from sfh_amd import mock, synth
    m, g = _synth_iso(9.0, -1.0)
    dmod, limit = 24.0, 3e5
    edges = synth.edges_200x300()
    imf = mock.Kroupa2001(0.08, 120.0)
    cat = sfh.generate_stars_mass(m, g, ["B", "V"], limit, imf, dist_mod=dmod, seed=31,
                                  binary_model=sfh.NoBinaries(), catalog=True)
    obs = sfh.model_cmd(cat, [synth.photerr, synth.photerr], [synth.completeness, synth.completeness], seed=32)
    H = mock.bin_cmd_catalog(obs, (0, 1), 1, edges).weights
    tpl = sfh.partial_cmd_smooth(m, [g[:, 0], g[:, 1]], [synth.photerr, synth.photerr], 1, (0, 1),
                                 synth.Kroupa2001(0.08, 120.0), [synth.completeness, synth.completeness], dmod=dmod,
                                 normalize_value=limit, mean_mass=imf.mean(), edges=edges).weights
    big = tpl >= 20.0
    chi2 = np.sum((H[big] - tpl[big]) ** 2 / tpl[big]) / big.sum()
    print(f"Monte Carlo: {H.sum():.0f} stars observed vs template {tpl.sum():.1f}; chi2/dof {chi2:.3f} over "
          f"{big.sum()} bins with >= 20 expected")
    assert abs(H.sum() - tpl.sum()) < 5 * math.sqrt(tpl.sum()) + 0.01 * tpl.sum()
    assert chi2 < 1.5
    cat.close()
    obs.close()
