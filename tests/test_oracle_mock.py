"""Pins of the oracle's mock-observation restatement (oracle/mock_ref.py) and
of the host code around the GPU path (mock.py): Philox4x64-10 against
NumPy's own implementation, the IMF sampler against its CDF, the
fixed-point stopping rule against the reference's sequential loop, the
reference's statistical checks of generate_stars_mass / generate_stars_mag
(test/cmd_simulation_test.jl) on the reference's test isochrone, model_cmd in
its closed-form limits, bin_cmd against numpy.histogram2d, and the
reference's host-side known answers (ingest_mags, sort_ingested,
mass_limits)."""
import math
import os
import sys

import numpy as np
import pytest
from scipy import integrate

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mock_ref as M   # noqa: E402

# the reference's test isochrone (test/cmd_simulation_test.jl:120-125)
M_INI = np.round(np.arange(1, 9) / 10.0, 10)
F606W = np.array([12.73830341586291, 10.017833567325411, 9.041398536300997, 8.331313172834694, 7.3892413746605765,
                  6.227971374447669, 4.93799980882831, -3.026])
F814W = np.array([11.137576129331675, 8.951605387408511, 8.064408199510426, 7.424658866406447, 6.6025729402403,
                  5.584714238650148, 4.448999828801471, -4.107])
NAMES = ["F606W", "F814W"]


def _salpeter():
    return M.PowerLaw([0.1, 100.0], [2.35], [1.0])


def _mass_frac(imf):
    # fraction of the birth mass in stars between min(m_ini) and max(m_ini)
    f = lambda m: m * m ** -2.35
    return integrate.quad(f, M_INI[0], M_INI[-1])[0] / integrate.quad(f, 0.1, 100.0)[0]


@pytest.mark.parametrize("key", [(0, 0), (12345, 678), (2 ** 64 - 1, 7)])
def test_philox_matches_numpy(key):
    g = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.zeros(4, dtype=np.uint64))
    ref = g.random_raw(32).reshape(8, 4)                     # numpy draws counters 1, 2, ...
    got = np.stack(M.philox4x64(np.arange(1, 9, dtype=np.uint64), 0, *key), axis=1)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("imf", [M.PowerLaw([0.1, 100.0], [2.35], [1.0]),
                                 M.PowerLaw([0.08, 0.5, 120.0], [1.3, 2.3], [1.0, 0.5]),
                                 M.PowerLaw([0.5, 2.0, 8.0], [1.0, 3.0], [1.0, 2.0])])
def test_imf_sampler_against_cdf(imf, sfh):
    # the CDF and mean of the restatement against quadrature of the pdf
    host = sfh.PowerLawIMF(imf.lo, 1 - imf.a, imf.coef)
    pdf = lambda m: float(host.pdf(np.array([m]))[0])
    for m in (imf.lo[0] * 1.5, float(np.sqrt(imf.lo[0] * imf.lo[-1])), imf.lo[-1] * 0.9):
        pts = [b for b in imf.lo[1:-1] if b < m]
        q = integrate.quad(pdf, imf.lo[0], m, points=pts or None, limit=200)[0]
        assert imf.cdf(m) == pytest.approx(q, rel=1e-9)
        assert float(host.cdf(m)[0]) == pytest.approx(q, rel=1e-9)
    mq = integrate.quad(lambda m: m * pdf(m), imf.lo[0], imf.lo[-1], points=list(imf.lo[1:-1]) or None,
                        limit=200)[0]
    assert imf.mean() == pytest.approx(mq, rel=1e-9) and host.mean() == pytest.approx(mq, rel=1e-9)
    # samples: Kolmogorov distance and mean
    n = 200_000
    u = M.u01(M.philox4x64(np.arange(n, dtype=np.uint64), 0, 99, 1)[1])
    s = np.sort(imf.sample(u))
    assert s[0] >= imf.lo[0] and s[-1] <= imf.lo[-1]
    F = imf.cdf(s)
    d = max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n))
    assert d < 1.63 / math.sqrt(n)                           # 1% level
    sd = math.sqrt(integrate.quad(lambda m: m * m * pdf(m), imf.lo[0], imf.lo[-1],
                                  points=list(imf.lo[1:-1]) or None, limit=200)[0] - mq * mq)
    assert abs(s.mean() - mq) < 5 * sd / math.sqrt(n)


def test_fixed_point_stop_rule_matches_sequential_loop():
    imf = _salpeter()
    i = np.arange(50_000, dtype=np.uint64)
    c = M.systems(i, imf, 1, 0.4, 0.0, 0.0, 5, 0).sum(axis=1)
    for limit in (1.0, 37.5, 1234.0, 8000.0):
        scale = 2.0 ** 62 / limit
        q = np.rint(np.minimum(c * scale, float(M.TARGET))).astype(np.uint64)
        k_fixed = int(np.flatnonzero(np.cumsum(q, dtype=np.uint64) >= np.uint64(M.TARGET))[0]) + 1
        assert k_fixed == M.sequential_stop(c, limit)


def test_generate_matches_sequential_definition():
    """Systems 0..K-1 with K from the sequential loop; kept = primary >= cut
    and positive flux (src/simulate.jl:322-336)."""
    imf = _salpeter()
    mini, mags = M_INI, np.stack([F606W, F814W], axis=1) + 20.0
    ms, g, idx, drawn = M.generate(mini, mags, 0, 500.0, imf, 1, 0.4, seed=3, stream=2, mmin_cut=0.15)
    allm = M.systems(np.arange(drawn + 10, dtype=np.uint64), imf, 1, 0.4, 0.0, 0.0, 3, 2)
    assert drawn == M.sequential_stop(allm.sum(axis=1), 500.0)
    flux = M.evaluate(allm[:drawn], mini, mags, 0)
    keep = ~(allm[:drawn, 0] < 0.15) & (flux.sum(axis=1) > 0)
    assert np.array_equal(idx, np.flatnonzero(keep))
    assert np.array_equal(ms, allm[:drawn][keep])
    # a system's magnitudes are those of its summed fluxes
    with np.errstate(divide="ignore"):
        assert np.allclose(g, -2.5 * np.log10(flux[keep]), rtol=0, atol=1e-12)


@pytest.mark.parametrize("bkind,factor", [(0, 1.0), (1, 1.4), (2, 1.0)])
def test_reference_total_mass_check(bkind, factor):
    # test/cmd_simulation_test.jl:133-166: total returned mass ~ total * mass_frac (x 1.4 with pairs)
    imf = _salpeter()
    mags = np.stack([F606W, F814W], axis=1) + 20.0
    ms, _, _, _ = M.generate(M_INI, mags, 0, 1e5, imf, bkind, 0.4, 0.1, 1.0, seed=11, stream=bkind,
                             mmin_cut=M_INI[0])
    assert ms.sum() == pytest.approx(1e5 * _mass_frac(imf) * factor, rel=5e-2)


@pytest.mark.parametrize("bkind", [0, 1, 2])
def test_reference_total_magnitude_check(bkind):
    # test/cmd_simulation_test.jl:171-205: -0.05 <= M_sampled - M_lim <= 0
    imf = _salpeter()
    dmod, absmag = 20.0, -8.0
    mags = np.stack([F606W, F814W], axis=1) + dmod
    limit = 10 ** (4 * (0 - (absmag + dmod)) / 10)
    _, g, _, _ = M.generate(M_INI, mags, 1, limit, imf, bkind, 0.4, 0.1, 1.0, limit_filter=1, seed=12,
                            stream=bkind)
    apparent = -2.5 * np.log10(np.sum(10 ** (-0.4 * g[:, 1])))
    assert -0.05 <= apparent - dmod - absmag <= 0.0


def test_model_cmd_limits():
    rng = np.random.default_rng(1)
    n = 100_000
    mags = rng.uniform(20, 26, (n, 2))
    tab = lambda v: (18.0, 0.01, np.full(1201, v))
    out, idx = M.model_cmd(mags, [tab(0.0)] * 2, [tab(1.0)] * 2, seed=4)
    assert np.array_equal(idx, np.arange(n)) and np.array_equal(out, mags)
    out, idx = M.model_cmd(mags, [tab(0.0)] * 2, [tab(0.6), tab(0.5)], seed=4)
    p = 0.3
    assert abs(idx.size / n - p) < 5 * math.sqrt(p * (1 - p) / n)
    out, idx = M.model_cmd(mags, [tab(0.1), tab(0.2)], [tab(1.0)] * 2, [tab(0.05), tab(0.0)], seed=4)
    z = (out - mags) - np.array([0.05, 0.0])
    for f, s in enumerate((0.1, 0.2)):
        assert abs(z[:, f].mean()) < 5 * s / math.sqrt(n)
        assert abs(z[:, f].std() / s - 1) < 5 / math.sqrt(2 * n)
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 5 / math.sqrt(n)


def test_table_eval_exact_on_linear_functions():
    x = np.linspace(17.5, 31.0, 1001)
    first, step = 18.0, 0.0125
    vals = 2.0 * (first + step * np.arange(801)) - 3.0
    got = M.table_eval(first, step, vals, x)
    ref = np.clip(2.0 * x - 3.0, vals[0], vals[-1])
    assert np.max(np.abs(got - ref)) < 1e-12


def test_bin_cmd_against_histogram2d():
    rng = np.random.default_rng(2)
    x, y, w = rng.uniform(-1, 3, 50_000), rng.uniform(17, 31, 50_000), rng.uniform(0, 2, 50_000)
    x0, dx, nx, y0, dy, ny = -0.5, 0.015, 200, 18.0, 0.04, 300
    ex, ey = x0 + dx * np.arange(nx + 1), y0 + dy * np.arange(ny + 1)
    h = M.bin_cmd(x, y, x0, dx, nx, y0, dy, ny)
    ref, _, _ = np.histogram2d(x, y, bins=[ex, ey])
    assert np.array_equal(h, ref)
    hw = M.bin_cmd(x, y, x0, dx, nx, y0, dy, ny, weights=w)
    refw, _, _ = np.histogram2d(x, y, bins=[ex, ey], weights=w)
    assert np.allclose(hw, refw, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ host code

def test_ingest_mags_reference_cases(sfh):
    # test/cmd_simulation_test.jl:15-40
    m = [1.0, 2.0, 3.0, 4.0]
    r1 = sfh.ingest_mags(m, np.array([[1, 1], [2, 2], [3, 3], [4, 4]], dtype=float))
    r2 = sfh.ingest_mags(m, np.array([[1, 2, 3, 4], [1, 2, 3, 4]], dtype=float))
    r3 = sfh.ingest_mags(m, [[1.0, 1.0], [2.0, 2.0], [3.0, 3.0], [4.0, 4.0]])
    r4 = sfh.ingest_mags(m, [[1.0, 2.0, 3.0, 4.0], [1.0, 2.0, 3.0, 4.0]])
    assert r1.shape == (4, 2)
    for r in (r2, r3, r4):
        assert np.array_equal(r, r1)
    for bad in (np.zeros((3, 3)), [[1.0, 1.0], [2.0, 2.0], [3.0, 3.0]], [[1.0] * 5, [1.0] * 5]):
        with pytest.raises(sfh.ArgumentError):
            sfh.ingest_mags(m, bad)
    with pytest.raises(sfh.ArgumentError):
        sfh.ingest_mags((1, 2, 3, 4), "not magnitudes")


def test_sort_ingested_reference_cases(sfh):
    # test/cmd_simulation_test.jl:42-62
    g = sfh.ingest_mags([1, 2, 3, 4], [[1.0, 2, 3, 4], [1.0, 2, 3, 4]])
    m2, g2 = sfh.sort_ingested([1.0, 2, 3, 4], g)
    assert np.array_equal(m2, [1, 2, 3, 4]) and np.array_equal(g2, g)
    m2, g2 = sfh.sort_ingested([2.0, 3, 1, 4], g)
    assert np.array_equal(m2, [1, 2, 3, 4]) and np.array_equal(g2, g[np.argsort([2, 3, 1, 4])])
    # repeated masses move to the next float
    m3, _ = sfh.sort_ingested([1.0, 2.0, 2.0, 2.0, 3.0], np.zeros((5, 1)))
    assert m3[2] == math.nextafter(2.0, 3.0) and m3[3] == math.nextafter(m3[2], 3.0) and np.all(np.diff(m3) > 0)


def test_mass_limits_reference_cases(sfh):
    # src/simulate.jl:83-88 (doctest) and test/cmd_simulation_test.jl:64-82
    m = [0.05, 0.1, 0.2, 0.3]
    assert sfh.mass_limits(m, [[4.0], [3.0], [2.0], [1.0]], ["F090W"], 2.5, "F090W") == (0.15, 0.3)
    assert sfh.mass_limits(m, [[4.0, 3.0], [3.0, 2.0], [2.0, 1.0], [1.0, 0.0]], ["F090W", "F150W"], 2.5,
                           "F090W") == (0.15, 0.3)
    assert sfh.mass_limits(m, [[4.0], [3.0], [2.0], [1.0]], ["F090W"], math.inf, "F090W") == (0.05, 0.3)
    assert sfh.mass_limits(m, [[4.0], [3.0], [2.0], [1.0]], ["F090W"], 10.0, "F090W") == (0.05, 0.3)
    with pytest.raises(sfh.DomainError):
        sfh.mass_limits(m, [[4.0], [3.0], [2.0], [1.0]], ["F090W"], 0.0, "F090W")
    with pytest.raises(sfh.ArgumentError):
        sfh.mass_limits(m, [[4.0], [3.0], [2.0], [1.0]], ["F090W"], 2.5, "F150W")


def test_binary_models_and_fractions(sfh):
    # test/cmd_simulation_test.jl:83-99
    assert sfh.binary_system_fraction(sfh.NoBinaries()) == 0
    assert sfh.binary_system_fraction(sfh.RandomBinaryPairs(0.4)) == 0.4
    assert sfh.BinaryMassRatio(0.4).fraction == 0.4
    with pytest.raises(sfh.ArgumentError):
        sfh.BinaryMassRatio(1.4)
    with pytest.raises(sfh.ArgumentError):
        sfh.BinaryMassRatio(0.4, 0.5, 1.5)


def test_surviving_fractions(sfh):
    imf = sfh.Salpeter1955(0.1, 100.0)
    masses = [np.linspace(0.1, 0.8, 20001), np.linspace(0.1, 2.0, 20001)]   # dense: trapz -> the integral
    coeffs = [1.0, 3.0]
    cdf = lambda v: float(imf.cdf(v)[0])
    ref = (1.0 * (cdf(0.8) - cdf(0.1)) + 3.0 * (cdf(2.0) - cdf(0.1))) / 4.0
    assert sfh.surviving_fraction(masses, coeffs, imf) == pytest.approx(ref, rel=1e-14)
    assert sfh.recycling_fraction(masses, coeffs, imf) == pytest.approx(1 - ref, rel=1e-14)
    smf = sfh.surviving_mass_fraction(masses, coeffs, imf)
    q = lambda hi: integrate.quad(lambda v: v * float(imf.pdf(np.array([v]))[0]), 0.1, hi)[0]
    assert smf == pytest.approx((q(0.8) + 3 * q(2.0)) / 4 / imf.mean(), rel=1e-6)
    assert sfh.recycling_mass_fraction(masses, coeffs, imf) == pytest.approx(1 - smf, rel=1e-14)


def test_composite_argument_errors(sfh):
    imf = sfh.Salpeter1955(0.1, 100.0)
    with pytest.raises(sfh.ArgumentError):
        sfh.generate_stars_mass_composite([M_INI, M_INI, M_INI], [[F606W, F814W]] * 2, NAMES, 1e5, [0.5, 0.5], imf)
    with pytest.raises(sfh.ArgumentError):
        sfh.generate_stars_mag_composite([M_INI] * 2, [[F606W, F814W]] * 2, NAMES, -8.0, "F814W", [0.5, 0.5], imf,
                                         frac_type="mass")
