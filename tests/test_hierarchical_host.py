"""Host parts of the hierarchical models (src/fitting/hierarchical/): the
reference's doctest values for the model types and the metallicity scale,
and calculate_coeffs' invariants (test/fitting/mzr_test.jl:9-45)."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def H(sfh):
    return sfh.hierarchical


def test_gaussian_dispersion_doctests(H):
    # dispersion_models.jl:63-88
    d = H.GaussianDispersion(0.2)
    assert d(1.0, 1.2) == pytest.approx(math.exp(-0.5), rel=1e-15)
    assert np.allclose(d.gradient(1.0, 1.2), (3.0326532985631656, -3.0326532985631656), rtol=1e-14)
    assert d.nparams == 1 and d.transforms() == (1,)
    assert H.GaussianDispersion(0.2, (False,)).free_params() == (False,)
    assert d.update_params([0.3]).sigma == 0.3
    with pytest.raises(H.ArgumentError):
        H.GaussianDispersion(-0.2)


def test_powerlaw_mzr_doctests(H):
    # mzr.jl:231-252
    m = H.PowerLawMZR(1.0, -1, 6)
    assert m(1e7) == pytest.approx(0.0, abs=1e-15)
    assert np.allclose(m.gradient(1e8), (2.0, 1.0, 1 / 1e8 / math.log(10)), rtol=1e-14)
    assert m.transforms() == (1, 0)
    u = H.PowerLawMZR(1.0, -1, 7, (True, False)).update_params((2.0, -2))
    assert (u.alpha, u.MH0, u.logMstar0, u.free) == (2.0, -2, 7, (True, False))
    with pytest.raises(H.ArgumentError):
        H.PowerLawMZR(-1.0, -1)


def test_amr_constructors(H):
    # amr.jl:217-223, 305-312
    a = H.LinearAMR.from_constraints((-2.5, 13.7), (-1.0, 0.0), 13.7)
    b = H.LinearAMR.from_constraints((-1.0, 0.0), (-2.5, 13.7), 13.7)
    assert (a.alpha, a.beta) == pytest.approx((b.alpha, b.beta), rel=1e-15)
    assert a(np.log10(13.7e9)) == pytest.approx(-2.5, abs=1e-12)
    assert a(-np.inf) == pytest.approx(-1.0, abs=1e-12)
    la = H.LogarithmicAMR.from_constraints((-2.5, 13.7), (-1.0, 0.0), 13.7)
    lb = H.LogarithmicAMR.from_constraints((-1.0, 0.0), (-2.5, 13.7), 13.7)
    assert (la.alpha, la.beta) == pytest.approx((lb.alpha, lb.beta), rel=1e-15)
    assert la(np.log10(13.7e9)) == pytest.approx(-2.5, abs=1e-10)
    with pytest.raises(H.ArgumentError):
        H.LinearAMR.from_constraints((-1.0, 13.7), (-2.5, 0.0))


def test_metallicity_scale(H):
    for mh in (-2.5, -1.0, 0.0, 0.3):
        assert H.MH_from_Z(H.Z_from_MH(mh)) == pytest.approx(mh, abs=1e-12)
    z, h = 0.004, 1e-7
    fd = (H.MH_from_Z(z + h) - H.MH_from_Z(z - h)) / (2 * h)
    assert H.dMH_dZ(z) == pytest.approx(fd, rel=1e-7)
    assert math.isnan(H.MH_from_Z(0.9))          # X <= 0


@pytest.mark.parametrize("kind", ["mzr", "amr"])
def test_calculate_coeffs_invariants(H, kind):
    rng = np.random.default_rng(94823)
    model = H.PowerLawMZR(1.0, -1.0, 6.0) if kind == "mzr" else H.LinearAMR(0.1, -2.0)
    disp = H.GaussianDispersion(0.2)
    ula = np.arange(10.0, 7.95, -0.1)
    umh = np.arange(-2.5, 0.001, 0.1)
    logAge = np.repeat(ula, umh.size)
    MH = np.tile(umh, ula.size)
    mstars = rng.uniform(size=ula.size) * 1e6
    x = H.calculate_coeffs(model, disp, mstars, logAge, MH)
    assert x.size == logAge.size
    assert np.allclose(x.reshape(ula.size, umh.size).sum(axis=1), mstars, rtol=1e-12)
    # the order of logAge is accounted for (mzr_test.jl:36-43)
    perm = rng.permutation(ula.size)
    y = H.calculate_coeffs(model, disp, mstars[perm], np.repeat(ula[perm], umh.size), np.tile(umh, ula.size))
    yy = y.reshape(ula.size, umh.size)[np.argsort(-ula[perm], kind="stable")].ravel()
    assert np.allclose(x, yy, rtol=1e-12)


def test_calculate_coeffs_hand_computed(H):
    """Two ages x three metallicities, worked by hand from mzr.jl:58-84 /
    amr.jl:56-77 and GaussianDispersion (dispersion_models.jl:91): the MZR's
    cumulative mass runs from the OLDEST age bin, μ = MH0 + α (log10 M* - 6),
    the AMR's μ = β + α (T_max - 10^(logAge-9)), and each age's weights
    exp(-((x - μ)/σ)^2 / 2) are normalised to its stellar mass."""
    logAge = [9.0, 9.0, 9.0, 10.0, 10.0, 10.0]
    MH = [-1.0, -0.5, 0.0, -1.0, -0.5, 0.0]
    mstars = [2e6, 1e6]                       # per unique age, in first-appearance order
    sigma = 0.2

    def weights(mu, m):
        w = [math.exp(-((x - mu) / sigma) ** 2 / 2) for x in (-1.0, -0.5, 0.0)]
        return [m * wi / sum(w) for wi in w]

    # MZR: the 10.0 bin is oldest, so its cumulative mass is its own 1e6;
    # the 9.0 bin's is 1e6 + 2e6
    mu_old = -1.0 + 1.0 * (math.log10(1e6) - 6.0)        # = -1
    mu_young = -1.0 + 1.0 * (math.log10(3e6) - 6.0)      # = -0.5229
    expect = weights(mu_young, 2e6) + weights(mu_old, 1e6)
    got = H.calculate_coeffs(H.PowerLawMZR(1.0, -1.0, 6.0), H.GaussianDispersion(sigma), mstars, logAge, MH)
    assert np.allclose(got, expect, rtol=1e-14, atol=0)
    # AMR with α = 0.1, β = -2, T_max = 13.7: μ(9) = -2 + 0.1 (13.7 - 1), μ(10) = -2 + 0.1 (13.7 - 10)
    expect = weights(-2.0 + 0.1 * (13.7 - 1.0), 2e6) + weights(-2.0 + 0.1 * (13.7 - 10.0), 1e6)
    got = H.calculate_coeffs(H.LinearAMR(0.1, -2.0), H.GaussianDispersion(sigma), mstars, logAge, MH)
    assert np.allclose(got, expect, rtol=1e-14, atol=0)
    # the same ages listed oldest first: masses follow unique(logAge) order
    got2 = H.calculate_coeffs(H.PowerLawMZR(1.0, -1.0, 6.0), H.GaussianDispersion(sigma), [1e6, 2e6],
                              logAge[3:] + logAge[:3], MH[3:] + MH[:3])
    assert np.allclose(got2, weights(mu_old, 1e6) + weights(mu_young, 2e6), rtol=1e-14, atol=0)


def test_stack_models(sfh):
    """stack_models (src/fitting/utilities.jl:12-13): reduce(hcat, vec.(models)),
    each template column-major."""
    a = np.arange(6.0).reshape(2, 3)
    b = 10 + np.arange(6.0).reshape(2, 3)
    s = sfh.stack_models([a, b])
    assert s.shape == (6, 2) and s.flags.f_contiguous
    assert np.array_equal(s[:, 0], [0, 3, 1, 4, 2, 5]) and np.array_equal(s[:, 1], [10, 13, 11, 14, 12, 15])
