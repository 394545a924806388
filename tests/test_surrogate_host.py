"""Host logic of the surrogate fill: which grid points fall back to the exact
partial_cmd_smooth (no GPU needed)."""
import numpy as np


def _mask(sfh, dgrid, ia, im, thr=0.5):
    from sfh_amd.surrogate import fallback_mask
    return fallback_mask(np.asarray(dgrid, dtype=np.float64), np.asarray(ia), np.asarray(im), thr)


def test_two_dimensional_cells(sfh):
    d = np.full((6, 5), np.nan)
    ia, im = [0, 3, 5], [0, 2, 4]
    for a in ia:
        for b in im:
            d[a, b] = 0.1
    d[3, 2] = 0.9                     # one bad corner: the four cells around it
    bad = _mask(sfh, d, ia, im)
    assert bad.all()
    d[3, 2] = 0.1
    d[5, 4] = 0.7                     # the far corner: only its cell
    bad = _mask(sfh, d, ia, im)
    expect = np.zeros((6, 5), dtype=bool)
    expect[3:6, 2:5] = True
    assert np.array_equal(bad, expect)
    d[5, 4] = 0.5                     # at the threshold: not above it
    assert not _mask(sfh, d, ia, im).any()


def test_single_metallicity_axis(sfh):
    """One metallicity: the cells are the age segments between checked ages."""
    d = np.full((11, 1), np.nan)
    ia, im = [0, 5, 10], [0]
    d[ia, 0] = [0.1, 0.1, 2.0]
    bad = _mask(sfh, d, ia, im)
    assert np.array_equal(np.flatnonzero(bad[:, 0]), np.arange(5, 11))
    d[ia, 0] = [3.0, 0.1, 0.1]
    assert np.array_equal(np.flatnonzero(_mask(sfh, d, ia, im)[:, 0]), np.arange(0, 6))


def test_single_age_axis_and_single_point(sfh):
    d = np.full((1, 9), np.nan)
    ia, im = [0], [0, 4, 8]
    d[0, im] = [0.1, 0.8, 0.1]
    assert np.array_equal(np.flatnonzero(_mask(sfh, d, ia, im)[0]), np.arange(0, 9))
    one = np.array([[0.9]])
    assert _mask(sfh, one, [0], [0]).all()
    assert not _mask(sfh, np.array([[0.2]]), [0], [0]).any()


def test_lattice_includes_ends(sfh):
    from sfh_amd.surrogate import _lattice
    assert list(_lattice(11, 5)) == [0, 5, 10]
    assert list(_lattice(12, 5)) == [0, 5, 10, 11]
    assert list(_lattice(1, 5)) == [0]

