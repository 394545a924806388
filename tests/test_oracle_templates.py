"""Pins the oracle's template-deposition functions: mpmath goldens of the
reference's kernel formulas and of a full bin_cmd_smooth, the reference's own
known answers for the host preparation, and invariants."""
import math

import numpy as np
import pytest

from helpers import golden, max_norm_err


def test_kernel_values_mpmath(O):
    for case in golden("kernels.json"):
        grid = case["grid"]
        vals = np.array(case["values"])
        for a, i in enumerate(grid):
            for b, j in enumerate(grid):
                if case["kind"] == "covariant":
                    v = O.gaussian_psf_covariant(i, j, case["h"], case["h"], case["x0"], case["y0"],
                                                 case["sx"], case["sy"], case["c"], case["A"])
                else:
                    v = O.gaussian_int_general(i - case["x0"], j - case["y0"], case["h"], case["h"],
                                               case["sx"], case["sy"], case["A"])
                ref = vals[a, b]
                assert abs(v - ref) <= 1e-12 * np.max(np.abs(vals)) + 1e-300, (case["kind"], case["sx"], i, j)


def test_bin_cmd_smooth_mpmath(O):
    for case in golden("template.json"):
        ex = O.Range(*case["xedges"][:2], case["xedges"][2])
        ey = O.Range(*case["yedges"][:2], case["yedges"][2])
        mat, pairs = O.bin_cmd_smooth(case["colors"], case["mags"], case["color_err"], case["mag_err"],
                                      case["cov_mult"], case["weights"], ex, ey)
        ref = np.array(case["matrix_colmajor"]).reshape(mat.shape, order="F")
        assert max_norm_err(mat, ref) < 1e-12, case["cov_mult"]
        assert pairs > 0


def test_bin_cmd_smooth_threads_identical(O):
    case = golden("template.json")[0]
    ex = O.Range(*case["xedges"][:2], case["xedges"][2])
    ey = O.Range(*case["yedges"][:2], case["yedges"][2])
    args = (case["colors"], case["mags"], case["color_err"], case["mag_err"], case["cov_mult"], case["weights"], ex, ey)
    a, pa = O.bin_cmd_smooth(*args, nthreads=1)
    b, pb = O.bin_cmd_smooth(*args, nthreads=4)
    assert np.array_equal(a, b) and pa == pb


@pytest.mark.parametrize("cov_mult", [-1, 1, 0])
def test_interior_kernel_sums_to_amplitude(O, cov_mult):
    # A kernel well inside the grid integrates to ~A (minus the truncation of
    # its +-7.5 sigma_x / +-5 sigma_y (covariant) or +-5 sigma (asymmetric) cutout).
    ex, ey = O.edges_range(-1.0, 1.0, 201), O.edges_range(-1.0, 1.0, 201)
    s = 0.05
    mat, _ = O.bin_cmd_smooth([0.013], [-0.021], [s], [s], cov_mult, [3.0], ex, ey)
    total = mat.sum()
    assert total == pytest.approx(3.0, rel=2e-5 if cov_mult == 0 else 2e-3)


def test_translation_by_whole_pixels(O):
    ex, ey = O.edges_range(0.0, 4.0, 81), O.edges_range(0.0, 4.0, 81)
    a, _ = O.bin_cmd_smooth([1.5013], [1.7021], [0.08], [0.06], 1, [1.0], ex, ey)
    b, _ = O.bin_cmd_smooth([1.5013 + 10 * ex.step], [1.7021 + 5 * ey.step], [0.08], [0.06], 1, [1.0], ex, ey)
    assert np.allclose(np.roll(np.roll(a, 10, axis=0), 5, axis=1)[20:60, 20:60], b[20:60, 20:60], rtol=1e-9, atol=1e-14)


def test_linearity_in_weights(O):
    case = golden("template.json")[2]
    ex = O.Range(*case["xedges"][:2], case["xedges"][2])
    ey = O.Range(*case["yedges"][:2], case["yedges"][2])
    a, _ = O.bin_cmd_smooth(case["colors"], case["mags"], case["color_err"], case["mag_err"], 1, case["weights"], ex, ey)
    b, _ = O.bin_cmd_smooth(case["colors"], case["mags"], case["color_err"], case["mag_err"], 1,
                            2.5 * np.array(case["weights"]), ex, ey)
    assert np.allclose(b, 2.5 * a, rtol=1e-14, atol=0)


def test_length_one_footprint_deposits_nothing(O):
    # addstar! :401 -- a footprint clipped to one column deposits nothing
    ex, ey = O.edges_range(0.0, 1.0, 11), O.edges_range(0.0, 1.0, 11)
    mat, pairs = O.bin_cmd_smooth([1.04], [0.5], [0.001], [0.3], 1, [1.0], ex, ey)
    assert pairs == 0 and not mat.any()


# ---------------------------------------------------------- host preparation

def test_mini_spacing_reference(O):
    # test/templates/template_test.jl:10-24
    m = [0.08, 0.10, 0.12, 0.14, 0.16]
    c = [1.0, 0.99, 0.98, 0.97, 0.96]
    y = [13.545, 12.899, 12.355, 11.459, 10.947]
    r = O.mini_spacing(m, c, y, 0.1, False)
    assert len(r) > 5 and max(np.diff(r)) < 0.1
    r, sp = O.mini_spacing(m, c, y, 0.1, True)
    assert np.allclose(sp, np.diff(r))


def test_midpoints_reference(O):
    # template_test.jl:26-33
    assert np.allclose(O.midpoints(list(np.arange(0.5, 1.0001, 0.1))), np.arange(0.55, 0.951, 0.1))
    assert np.allclose(O.midpoints([1.0, 0.9, 0.8, 0.7, 0.6, 0.5]), [0.95, 0.85, 0.75, 0.65, 0.55])
    assert np.allclose(O.midpoints([1.0, 2.0, 2.2, 2.1]), [1.5, 2.1, 2.15])


def test_interpolate_mini_doctest(O):
    # src/StarFormationHistories.jl:55-66
    m = [0.08, 0.10, 0.12, 0.14, 0.16]
    mags = [13.545, 12.899, 12.355, 11.459, 10.947]
    new = [0.08 + 0.01 * k for k in range(9)]
    res = [13.545, 13.222, 12.899, 12.626999999999999, 12.355, 11.907, 11.459, 11.203, 10.947]
    assert np.allclose(O.interpolate_mini(m, mags, new), res, rtol=1e-12)


def test_callables_reference_values(O):
    # test/utilities/utilities_test.jl:44-47 (BigFloat references)
    assert O.martin2016_complete(20.0, 1.0, 25.0, 1.0) == pytest.approx(0.9933071490757151444, rel=1e-15)
    assert O.exp_photerr(20.0, 1.05, 10.0, 32.0, 0.01) == pytest.approx(0.01286605230281143891, rel=1e-14)


def test_calculate_weights_trapezoid(O):
    # with a constant pdf and unit completeness the weights are the mass steps
    w = O.calculate_weights([1.0, 1.5, 2.5], [1.0, 1.0, 1.0], lambda m: 2.0, 3.0, 6.0)
    assert np.allclose(w, [0.5 * 2 * 3 / 6, 1.0 * 2 * 3 / 6])
    # the completeness of the LEFT point is used (:773)
    w = O.calculate_weights([1.0, 2.0], [0.5, 1.0], lambda m: 1.0, 1.0, 1.0)
    assert w == [0.5]
