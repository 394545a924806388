"""Pins of the oracle's binary-system restatement (binary_hess,
src/StarFormationHistories.jl:627-706) and of the host helpers around it:
the reference's doctest values, the definitions of the binary fractions,
two independent restatements of the pair histogram against each other, and
the closed form of its total weight."""
import math

import numpy as np
import pytest


def test_flux_magnitude_doctests(O, sfh):
    # src/utilities.jl:53,63
    assert O.mag2flux(15.0, 25.0) == pytest.approx(10 ** (4 * (25.0 - 15.0) / 10), rel=1e-15)
    assert O.flux2mag(10000.0, 25.0) == pytest.approx(25.0 - 5 * math.log10(10000.0) / 2, rel=1e-15)
    assert sfh.mag2flux(15.0, 25.0) == pytest.approx(1e4, rel=1e-15)
    assert sfh.flux2mag(1e4, 25.0) == pytest.approx(15.0, rel=1e-15)
    # two equal stars are 2.5 log10(2) brighter than one
    assert O.flux2mag(2 * O.mag2flux(20.0)) == pytest.approx(20.0 - 2.5 * math.log10(2.0), abs=1e-13)


def test_binary_fractions(sfh):
    # src/simulate.jl:139-187
    m = sfh.RandomBinaryPairs(0.3)
    assert sfh.binary_system_fraction(m) == 0.3
    assert sfh.binary_number_fraction(m) == pytest.approx(0.6 / 1.3, rel=1e-15)
    assert sfh.binary_mass_fraction(m, None) == pytest.approx(0.6 / 1.3, rel=1e-15)
    assert sfh.binary_mass_fraction(sfh.NoBinaries(), None) == 0.0
    with pytest.raises(sfh.ArgumentError):
        sfh.RandomBinaryPairs(1.5)
    assert sfh.trapz([0.0, 1.0, 3.0], [1.0, 3.0, 2.0]) == pytest.approx(0.5 * 4 + 1.0 * 5)


def _tiny_isochrone(n, seed):
    rng = np.random.default_rng(seed)
    m = np.sort(rng.uniform(0.1, 1.5, n))
    ma = 28.0 - 7.0 * np.log10(m / 0.1) + 0.05 * rng.normal(size=n)
    mb = ma - 0.4 - 0.3 * np.log10(m / 0.1)
    imf = m ** -2.3
    return m, imf, ma, mb


def test_pair_histogram_loops_vs_vectorised(O):
    m, imf, ma, mb = _tiny_isochrone(30, 3)
    ex, ey = O.Range(-1.0, 0.05, 41), O.Range(18.0, 0.1, 121)
    compl = [lambda v: 1.0 / (1.0 + math.exp((v - 27.0) / 0.7)), lambda v: 1.0 / (1.0 + math.exp((v - 26.5) / 0.5))]
    bias = [lambda v: 0.01 * (v - 20.0), lambda v: -0.02]
    for y_is_a in (True, False):
        a = O.binary_pairs_hist_loops(m, imf, ma, mb, y_is_a, 0.37, compl, bias, ex, ey)
        b = O.binary_pairs_hist(m, imf, ma, mb, y_is_a, 0.37, compl, bias, ex, ey)
        assert a.sum() > 0
        assert np.allclose(a, b, rtol=1e-13, atol=1e-15 * a.max())


def test_pair_histogram_total_weight(O):
    """With completeness 1, bias 0 and edges that hold every pair, the
    histogram sums to prefac * sum_{i<j} a_i a_j, a_i = dM_i (imf_i + imf_i+1)
    = prefac * ((sum a)^2 - sum a^2) / 2."""
    m, imf, ma, mb = _tiny_isochrone(200, 5)
    ex, ey = O.Range(-3.0, 0.01, 601), O.Range(10.0, 0.05, 441)
    h = O.binary_pairs_hist(m, imf, ma, mb, True, 0.5, [lambda v: 1.0] * 2, [lambda v: 0.0] * 2, ex, ey)
    a = np.diff(m) * (imf[:-1] + imf[1:])
    assert h.sum() == pytest.approx(0.5 * (a.sum() ** 2 - (a ** 2).sum()) / 2, rel=1e-12)
