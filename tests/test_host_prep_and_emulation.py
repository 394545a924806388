"""CPU checks of the product's host code and of the deposition kernel's FP32
arithmetic (host build of the same __host__ __device__ functions, arranged like
the kernel) against the FP64 oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import golden, max_norm_err, oracle_range

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _emu(emu, O, colors, mags, cerr, merr, weights, cov_mult, ex, ey, cmult=None):
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    c, m, ce, me, w = map(d, (colors, mags, cerr, merr, weights))
    cm = d(cmult) if cmult is not None else None
    out = np.zeros((ex.len - 1) * (ey.len - 1))
    P = ctypes.POINTER(ctypes.c_double)
    pairs = emu.emu_bin_cmd_smooth(c.ctypes.data_as(P), m.ctypes.data_as(P), ce.ctypes.data_as(P),
                                   me.ctypes.data_as(P), w.ctypes.data_as(P),
                                   cm.ctypes.data_as(P) if cm is not None else None, c.size, float(cov_mult),
                                   ex, ey, out.ctypes.data_as(P))
    return out.reshape((ex.len - 1, ey.len - 1), order="F"), pairs


def test_emulation_matches_golden(emu, O):
    for case in golden("template.json"):
        ex = O.Range(*case["xedges"][:2], case["xedges"][2])
        ey = O.Range(*case["yedges"][:2], case["yedges"][2])
        mat, _ = _emu(emu, O, case["colors"], case["mags"], case["color_err"], case["mag_err"],
                      case["weights"], case["cov_mult"], ex, ey)
        ref = np.array(case["matrix_colmajor"]).reshape(mat.shape, order="F")
        assert max_norm_err(mat, ref) < 2e-6, case["cov_mult"]


@pytest.mark.parametrize("cov_mult", [1, -1, 0])
def test_emulation_synthetic_template(sfh, emu, O, cov_mult):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    sp, _ = synth.template_points(synth.TemplateSpec(9.3, -1.2, 42), 500, edges, 2.0)
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    ref, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, cov_mult, sp.weights, ex, ey, nthreads=8)
    mat, pe = _emu(emu, O, sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.weights, cov_mult, ex, ey)
    assert pe == pr
    assert max_norm_err(mat, ref) < 1e-5


def test_emulation_real_valued_covariance(sfh, emu, O):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    sp, cm = synth.template_points(synth.TemplateSpec(8.4, -0.3, 5), 500, edges, 2.0, correlation=0.6)
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    ref, _ = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey, cmult=cm, nthreads=8)
    mat, _ = _emu(emu, O, sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.weights, 1, ex, ey, cmult=cm)
    assert max_norm_err(mat, ref) < 1e-5


@pytest.mark.parametrize("la,mh,cov_mult", [(6.6, -2.5, 1), (10.05, 0.0, 1), (8.0, -1.0, 0)])
def test_emulation_config1_templates(sfh, emu, O, la, mh, cov_mult):
    """The kernel arithmetic at the 700-template workload's shape: a 200 x 300
    grid, 2000-point isochrones, IMF 0.08-120 Msun, colour-magnitude
    correlation 0.6 (covariant) or the pixel-space asymmetric kernel; every
    x form (the Gaussian form, the degree-3 series, erfc differences) occurs.  Max-normalised
    error <= 1e-5 and per-bin relative error <= 2e-5 on bins above 1e-6 of the
    maximum."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    sp, cm = synth.template_points(synth.TemplateSpec(la, mh, 7), 2000, edges, 120.0, correlation=0.6)
    if cov_mult == 0:
        cm = None
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    ref, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, cov_mult, sp.weights, ex, ey,
                               cmult=cm, nthreads=8)
    mat, pe = _emu(emu, O, sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.weights, cov_mult, ex, ey, cmult=cm)
    assert pe == pr
    assert max_norm_err(mat, ref) <= 1e-5
    big = ref > 1e-6 * ref.max()
    assert np.max(np.abs(mat[big] - ref[big]) / ref[big]) <= 2e-5


def test_host_prep_matches_oracle_prep(sfh, O):
    """The product's vectorised partial_cmd_smooth preparation equals the
    oracle's loop restatement."""
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    m, mags = synth.isochrone(9.0, -1.0, 500, 17, mmax_imf=2.0)
    for y_index, cidx in ((1, (0, 1)), (0, (0, 1))):
        a = sfh.partial_cmd_smooth_points(m, mags, [synth.photerr, synth.photerr], y_index, cidx, imf,
                                          [synth.completeness, synth.completeness], dmod=24.0,
                                          normalize_value=1e3, mean_mass=imf.mean(), edges=edges)
        b = O.partial_cmd_smooth_points(m, mags, [lambda v: float(synth.photerr(v))] * 2, y_index, cidx,
                                        lambda v: float(imf(v)),
                                        [lambda v: float(synth.completeness(v))] * 2, dmod=24.0,
                                        normalize_value=1e3, mean_mass=imf.mean(),
                                        ex=oracle_range(O, edges[0]), ey=oracle_range(O, edges[1]))
        assert a.cov_mult == b.cov_mult
        for f in ("colors", "mags", "color_err", "mag_err", "weights"):
            assert np.allclose(getattr(a, f), getattr(b, f), rtol=1e-13, atol=0), f


def test_mini_spacing_product_equals_oracle(sfh, O):
    rng = np.random.default_rng(9)
    m = np.sort(rng.uniform(0.1, 2.0, 200))
    c = np.cumsum(rng.normal(0, 0.05, 200))
    y = np.cumsum(rng.normal(-0.05, 0.1, 200))
    a = sfh.mini_spacing(m, c, y, 0.03)
    b = O.mini_spacing(m, c, y, 0.03)
    assert np.array_equal(a, np.array(b))


def test_kroupa_imf_normalised(sfh):
    from scipy.integrate import quad
    from sfh_amd import synth
    imf = synth.Kroupa2001(0.08, 120.0)
    assert quad(imf, 0.08, 0.5)[0] + quad(imf, 0.5, 120.0, limit=200)[0] == pytest.approx(1.0, rel=1e-9)
    mm = quad(lambda x: x * imf(x), 0.08, 0.5)[0] + quad(lambda x: x * imf(x), 0.5, 120.0, limit=200)[0]
    assert imf.mean() == pytest.approx(mm, rel=1e-9)


# ------------------------------------------------------------------ the ABI

def _header_symbols():
    with open(os.path.join(ROOT, "include", "sfh.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|const float\*)\s+(sfh_\w+)\s*\(", text, re.M)))


def test_library_exports_header(sfh):
    from sfh_amd import _lib
    L = _lib.load()
    syms = _header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(syms) == bound


def test_errors_without_gpu_are_loud(sfh):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(Exception):
        sfh.Context(0)


def test_isochrone_batch_packing(sfh):
    """The packed isochrone form: concatenation order, offsets and the
    argument checks (no GPU needed)."""
    a = (np.array([0.1, 0.2, 0.3]), [np.array([1.0, 2.0, 3.0]), np.array([4.0, 5.0, 6.0])])
    b = (np.array([0.5, 0.6]), [np.array([7.0, 8.0]), np.array([9.0, 10.0])])
    ib = sfh.IsochroneBatch.from_list([a, b])
    assert len(ib) == 2 and ib.nfilters == 2
    assert list(ib.offsets) == [0, 3, 5]
    assert list(ib.m_ini) == [0.1, 0.2, 0.3, 0.5, 0.6]
    assert list(ib.mags[1]) == [4.0, 5.0, 6.0, 9.0, 10.0]
    with pytest.raises(sfh.ArgumentError):
        sfh.IsochroneBatch(np.zeros(3), [np.zeros(3)], [0, 2])          # offsets must end at length(m_ini)
    with pytest.raises(sfh.ArgumentError):
        sfh.IsochroneBatch(np.zeros(3), [np.zeros(2)], [0, 3])          # filter length
    with pytest.raises(sfh.ArgumentError):
        sfh.IsochroneBatch(np.zeros(3), [np.zeros(3)], [0, 2, 1, 3])    # decreasing offsets
