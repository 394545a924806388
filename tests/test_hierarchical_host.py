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
