"""CPU FP64 oracle for the StarFormationHistories.jl hot paths.

TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module.  The product package never imports it and has no CPU
fallback.

Two layers:

* ``sfh_oracle.c`` (built to ``oracle/build/libsfh_oracle.so``) restates the
  numerical kernels: composite!/loglikelihood/∇loglikelihood!/fg! and
  addstar!/bin_cmd_smooth with both PSF kernels.
* This module restates the host-side template preparation of
  ``partial_cmd_smooth`` with plain Python loops (inputs are a few thousand
  points): mini_spacing, interpolate_mini, midpoints, calculate_weights, the
  photometric-error and completeness callables.

Citations are into the StarFormationHistories.jl v1.3.1 source tree.
"""
from __future__ import annotations

import bisect
import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libsfh_oracle.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class Range(C.Structure):
    """A StepRangeLen of bin edges: first value, step, number of edges."""
    _fields_ = [("first", C.c_double), ("step", C.c_double), ("len", C.c_int64)]


def edges_range(start: float, stop: float, length: int) -> Range:
    """``range(start, stop; length)`` as used by calculate_edges
    (src/StarFormationHistories.jl:472-473).  Julia's Float64 range keeps its
    step in twice precision and rounds it once; (stop-start)/(length-1) agrees
    to within 1 ulp, which no tolerance in the parity tests resolves."""
    return Range(float(start), (float(stop) - float(start)) / (length - 1), int(length))


def range_values(r: Range) -> np.ndarray:
    return r.first + np.arange(r.len, dtype=np.float64) * r.step


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_composite_f32.argtypes = [_fp, _dp, C.c_int64, C.c_int64, _dp, C.c_int]
        L.orc_composite_f64.argtypes = [_dp, _dp, C.c_int64, C.c_int64, _dp, C.c_int]
        L.orc_loglikelihood.argtypes = [_dp, _dp, C.c_int64, C.c_double]
        L.orc_loglikelihood.restype = C.c_double
        L.orc_grad_single_f64.argtypes = [_dp, _dp, _dp, C.c_int64, C.c_double]
        L.orc_grad_single_f64.restype = C.c_double
        L.orc_grad_inplace_f32.argtypes = [_dp, _dp, _fp, _dp, C.c_int64, C.c_int64, C.c_double, C.c_int]
        L.orc_grad_inplace_f64.argtypes = [_dp, _dp, _dp, _dp, C.c_int64, C.c_int64, C.c_double, C.c_int]
        L.orc_fg_f32.argtypes = [C.c_int, _dp, _dp, _fp, _dp, _dp, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                 C.c_int]
        L.orc_fg_f32.restype = C.c_double
        L.orc_fg_batched_f32.argtypes = [_fp, _dp, _dp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                         _dp, _dp, C.c_int]
        L.orc_histogram_pix.argtypes = [C.c_double, Range]
        L.orc_histogram_pix.restype = C.c_double
        L.orc_histogram_data.argtypes = [C.c_double, Range]
        L.orc_histogram_data.restype = C.c_double
        L.orc_gaussian_psf_covariant.argtypes = [C.c_double] * 11
        L.orc_gaussian_psf_covariant.restype = C.c_double
        L.orc_gaussian_int_general.argtypes = [C.c_double] * 8
        L.orc_gaussian_int_general.restype = C.c_double
        L.orc_addstar_covariant.argtypes = [_dp, Range, Range] + [C.c_double] * 6
        L.orc_addstar_covariant.restype = C.c_int64
        L.orc_addstar_asymmetric.argtypes = [_dp, C.c_int64, C.c_int64] + [C.c_double] * 5
        L.orc_addstar_asymmetric.restype = C.c_int64
        L.orc_bin_cmd_smooth.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_double,
                                         Range, Range, _dp, C.c_int]
        L.orc_bin_cmd_smooth.restype = C.c_int64
        L.orc_count_pairs.argtypes = [_dp, _dp, _dp, _dp, C.c_int64, C.c_double, Range, Range]
        L.orc_count_pairs.restype = C.c_int64
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def max_threads() -> int:
    return int(lib().orc_max_threads())


# --------------------------------------------------------------------------
# Template-mixture likelihood (src/fitting/fitting_base.jl, solvers.jl)
# --------------------------------------------------------------------------

def _models_colmajor(models):
    """models: (N, T) array; returns a Fortran-ordered view/copy and dtype tag."""
    m = np.asfortranarray(models)
    if m.dtype == np.float32:
        return m, "f32"
    return np.asfortranarray(m, dtype=np.float64), "f64"


def composite(models, coeffs, nthreads: int = 1) -> np.ndarray:
    """composite! (src/fitting/fitting_base.jl:55-65)."""
    m, tag = _models_colmajor(models)
    N, T = m.shape
    c = _f64(coeffs)
    assert c.shape == (T,)
    out = np.empty(N, dtype=np.float64)
    if tag == "f32":
        lib().orc_composite_f32(_f(m), _d(c), N, T, _d(out), nthreads)
    else:
        lib().orc_composite_f64(_d(m), _d(c), N, T, _d(out), nthreads)
    return out


def loglikelihood(comp, data, eps: float = np.finfo(np.float64).eps) -> float:
    """loglikelihood(composite, data) (src/fitting/fitting_base.jl:84-96)."""
    cm, dt = _f64(comp).ravel(), _f64(data).ravel()
    assert cm.shape == dt.shape
    return float(lib().orc_loglikelihood(_d(cm), _d(dt), cm.size, eps))


def grad_single(model, comp, data, eps: float = np.finfo(np.float64).eps) -> float:
    """∇loglikelihood(model, composite, data) (src/fitting/fitting_base.jl:144-160)."""
    md, cm, dt = _f64(model).ravel(), _f64(comp).ravel(), _f64(data).ravel()
    return float(lib().orc_grad_single_f64(_d(md), _d(cm), _d(dt), md.size, eps))


def grad_inplace(G, comp, models, data, eps: float = np.finfo(np.float64).eps, nthreads: int = 1):
    """∇loglikelihood!(G, composite, models, data) (src/fitting/fitting_base.jl:265-285).
    G and comp are float64 arrays modified in place (comp becomes 1 - n/c)."""
    m, tag = _models_colmajor(models)
    N, T = m.shape
    dt = _f64(data)
    assert G.dtype == np.float64 and comp.dtype == np.float64 and G.flags.c_contiguous and comp.flags.c_contiguous
    if tag == "f32":
        lib().orc_grad_inplace_f32(_d(G), _d(comp), _f(m), _d(dt), N, T, eps, nthreads)
    else:
        lib().orc_grad_inplace_f64(_d(G), _d(comp), _d(m), _d(dt), N, T, eps, nthreads)
    return G


def fg(models, data, coeffs, want_F=True, want_G=True, eps=np.finfo(np.float64).eps, nthreads=1,
       eps_grad=None):
    """fg!(F, G, coeffs, models, data, composite) (src/fitting/solvers.jl:20-38),
    for float32 models.  Returns (negLL or None, G or None, composite-after).
    ``eps`` clamps the log-likelihood (eps of the promoted type,
    fitting_base.jl:86), ``eps_grad`` the gradient's R (eps(eltype(composite)),
    :277); it defaults to ``eps``."""
    eps_grad = eps if eps_grad is None else eps_grad
    m = np.asfortranarray(models, dtype=np.float32)
    N, T = m.shape
    comp = np.empty(N, dtype=np.float64)
    G = np.empty(T, dtype=np.float64) if want_G else None
    c = _f64(coeffs)
    dt = _f64(data)
    r = lib().orc_fg_f32(1 if want_F else 0, _d(G) if G is not None else None, _d(c), _f(m),
                         _d(dt), _d(comp), N, T, eps, eps_grad, nthreads)
    return (float(r) if want_F else None), G, comp


def fg_batched(models, data, X, eps=np.finfo(np.float64).eps, want_G=True, nthreads=1, eps_grad=None):
    """B independent fg! calls, one per column of X (T x B)."""
    eps_grad = eps if eps_grad is None else eps_grad
    m = np.asfortranarray(models, dtype=np.float32)
    N, T = m.shape
    Xf = np.asfortranarray(X, dtype=np.float64)
    assert Xf.shape[0] == T
    B = Xf.shape[1]
    negll = np.empty(B, dtype=np.float64)
    G = np.empty((T, B), dtype=np.float64, order="F") if want_G else None
    lib().orc_fg_batched_f32(_f(m), _d(_f64(data)), _d(Xf), B, N, T, eps, eps_grad, _d(negll),
                             _d(G) if G is not None else None, nthreads)
    return negll, G


# --------------------------------------------------------------------------
# Template deposition kernels (src/StarFormationHistories.jl)
# --------------------------------------------------------------------------

def histogram_pix(d, e: Range) -> float:
    return float(lib().orc_histogram_pix(d, e))


def histogram_data(x, e: Range) -> float:
    return float(lib().orc_histogram_data(x, e))


def gaussian_psf_covariant(x, y, hx, hy, x0, y0, sx, sy, c, A, B=0.0) -> float:
    return float(lib().orc_gaussian_psf_covariant(x, y, hx, hy, x0, y0, sx, sy, c, A, B))


def gaussian_int_general(dx, dy, hx, hy, sx, sy, A, B=0.0) -> float:
    return float(lib().orc_gaussian_int_general(dx, dy, hx, hy, sx, sy, A, B))


def bin_cmd_smooth(colors, mags, color_err, mag_err, cov_mult, weights, ex: Range, ey: Range,
                   cmult=None, nthreads: int = 1):
    """bin_cmd_smooth (src/StarFormationHistories.jl:574-621).  Returns
    (matrix (nx, ny) float64 in Fortran order, pair count)."""
    c, m, ce, me, w = (_f64(a) for a in (colors, mags, color_err, mag_err, weights))
    P = c.size
    assert all(a.size == P for a in (m, ce, me, w))
    nx, ny = ex.len - 1, ey.len - 1
    out = np.zeros(nx * ny, dtype=np.float64)
    cm = _f64(cmult) if cmult is not None else None
    pairs = lib().orc_bin_cmd_smooth(_d(c), _d(m), _d(ce), _d(me), _d(w),
                                     _d(cm) if cm is not None else None, P, float(cov_mult),
                                     ex, ey, _d(out), nthreads)
    return out.reshape((nx, ny), order="F"), int(pairs)


def count_pairs(colors, mags, color_err, mag_err, cov_mult, ex: Range, ey: Range) -> int:
    c, m, ce, me = (_f64(a) for a in (colors, mags, color_err, mag_err))
    return int(lib().orc_count_pairs(_d(c), _d(m), _d(ce), _d(me), c.size, float(cov_mult), ex, ey))


# --------------------------------------------------------------------------
# Host-side template preparation (pure-Python restatement)
# --------------------------------------------------------------------------

def martin2016_complete(m, A, m50, rho):
    """Martin2016_complete (src/utilities.jl:205)."""
    return A / (1.0 + math.exp((m - m50) / rho))


def exp_photerr(m, a, b, c, d):
    """exp_photerr (src/utilities.jl:218)."""
    return a ** (b * (m - c)) + d


def mini_spacing(m_ini, colors, mags, dmag, ret_spacing=False):
    """mini_spacing (src/StarFormationHistories.jl:110-148)."""
    m_ini, colors, mags = list(map(float, m_ini)), list(map(float, colors)), list(map(float, mags))
    assert len(m_ini) == len(colors) == len(mags)
    new_mini = [m_ini[0]]                       # :116-117, before sorting
    idx = sorted(range(len(m_ini)), key=lambda k: m_ini[k])   # :119 (stable sortperm)
    m_ini = [m_ini[k] for k in idx]
    mags = [mags[k] for k in idx]
    colors = [colors[k] for k in idx]
    for i in range(len(m_ini) - 1):             # :125
        dx = colors[i + 1] - colors[i]
        dy = mags[i + 1] - mags[i]
        diffi = math.hypot(dx, dy)
        if diffi > dmag:
            npoints = math.ceil(diffi / dmag)   # round(Int, ·, RoundUp)
            mass_step = (m_ini[i + 1] - m_ini[i]) / npoints
            for j in range(1, npoints + 1):
                new_mini.append(m_ini[i] + mass_step * j)
        else:
            new_mini.append(m_ini[i + 1])
    seen, uniq = set(), []                      # :141 unique, first occurrence kept
    for v in new_mini:
        if v not in seen:
            seen.add(v)
            uniq.append(v)
    if not ret_spacing:
        return uniq
    return uniq, [uniq[k + 1] - uniq[k] for k in range(len(uniq) - 1)]


def interpolate_mini(m_ini, mags, new_mini):
    """interpolate_mini (src/StarFormationHistories.jl:77-81) with sort_ingested
    (src/simulate.jl:54-63) and Interpolations' Gridded(Linear()).  Knots must be
    distinct (duplicate knots would be moved by deduplicate_knots!, which this
    restatement does not reproduce) and queries must lie inside the knots
    (Interpolations throws BoundsError otherwise)."""
    order = sorted(range(len(m_ini)), key=lambda k: float(m_ini[k]))
    k = [float(m_ini[i]) for i in order]
    y = [float(mags[i]) for i in order]
    for a, b in zip(k, k[1:]):
        if not a < b:
            raise ValueError("duplicate isochrone masses: deduplicate_knots! semantics not restated")
    out = []
    for x in new_mini:
        x = float(x)
        if x < k[0] or x > k[-1]:
            raise ValueError(f"BoundsError: {x} outside [{k[0]}, {k[-1]}]")
        i = bisect.bisect_right(k, x) - 1
        i = min(max(i, 0), len(k) - 2)
        t = (x - k[i]) / (k[i + 1] - k[i])
        out.append(y[i] * (1.0 - t) + y[i + 1] * t)
    return out


_RTOL_DEFAULT = math.sqrt(np.finfo(np.float64).eps)


def _isapprox(x, y):
    return x == y or (math.isfinite(x) and math.isfinite(y) and
                      abs(x - y) <= _RTOL_DEFAULT * max(abs(x), abs(y)))


def midpoints(v, ranges=True):
    """midpoints(v::AbstractVector, ranges=true) (src/StarFormationHistories.jl:417-431)."""
    v = [float(a) for a in v]
    d = [v[i + 1] - v[i] for i in range(len(v) - 1)]
    if ranges and all(_isapprox(d[0], di) for di in d):
        vstep = d[0]
        start = v[0] + vstep / 2
        stop = v[-1] - vstep / 2
        n = len(v) - 1
        if n == 1:
            return [start]
        st = (stop - start) / (n - 1)
        return [start + i * st for i in range(n)]
    return [v[i] + d[i] / 2 for i in range(len(d))]


def calculate_weights(mini, completeness, imf, normalize_value, mean_mass, spacing=None):
    """calculate_weights (src/StarFormationHistories.jl:765-776): trapezoid of the
    IMF pdf over each mass step, times the completeness of the LEFT point."""
    if spacing is None:
        spacing = [mini[i + 1] - mini[i] for i in range(len(mini) - 1)]
    assert len(mini) == len(completeness) and len(spacing) == len(mini) - 1
    w = []
    for i in range(len(mini) - 1):
        wi = spacing[i] * (imf(mini[i]) + imf(mini[i + 1])) / 2
        wi *= completeness[i] * normalize_value / mean_mass
        w.append(wi)
    return w


@dataclass
class SmoothPoints:
    """The arrays partial_cmd_smooth hands to bin_cmd_smooth (:886-888), and
    the resampled isochrone its binary branch uses."""
    colors: np.ndarray
    mags: np.ndarray
    color_err: np.ndarray
    mag_err: np.ndarray
    weights: np.ndarray
    cov_mult: int
    new_mini: list = None
    new_iso: list = None


def partial_cmd_smooth_points(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                              completeness_funcs=None, bias_funcs=None, *, dmod=0.0,
                              normalize_value=1.0, mean_mass, ex: Range, ey: Range) -> SmoothPoints:
    """Host preparation of partial_cmd_smooth (src/StarFormationHistories.jl:829-888),
    NoBinaries only.  Indices are 0-based here (y_index, color_indices)."""
    nf = len(mags)
    if completeness_funcs is None:
        completeness_funcs = [lambda m: 1.0] * nf
    if bias_funcs is None:
        bias_funcs = [lambda m: 0.0] * nf
    assert len(color_indices) == 2
    assert nf == len(mag_err_funcs) == len(completeness_funcs) == len(bias_funcs)
    c0, c1 = color_indices
    ymags = [float(a) for a in mags[y_index]]
    colors = [float(a) - float(b) for a, b in zip(mags[c0], mags[c1])]        # :849
    dmag = min(ex.step, ey.step)                                               # :851
    new_mini, new_spacing = mini_spacing(m_ini, colors, ymags, dmag, True)     # :852
    new_iso = [[v + dmod for v in interpolate_mini(m_ini, mg, new_mini)] for mg in mags]  # :854
    mag_err = [[mag_err_funcs[i](v) for v in new_iso[i]] for i in range(nf)]  # :855
    if y_index in color_indices:                                               # :856
        x_c_idx = c1 if c0 == y_index else c0
        color_err = mag_err[x_c_idx]
        compl = [completeness_funcs[c0](a) * completeness_funcs[c1](b)
                 for a, b in zip(new_iso[c0], new_iso[c1])]
        weights = calculate_weights(new_mini, compl, imf, normalize_value, mean_mass, new_spacing)
        cov_mult = -1 if y_index == c0 else 1                                  # :868
    else:
        color_err = [math.sqrt(a * a + b * b) for a, b in zip(mag_err[c0], mag_err[c1])]  # :871
        compl = [completeness_funcs[c0](a) * completeness_funcs[c1](b) * completeness_funcs[y_index](c)
                 for a, b, c in zip(new_iso[c0], new_iso[c1], new_iso[y_index])]
        weights = calculate_weights(new_mini, compl, imf, normalize_value, mean_mass, new_spacing)
        cov_mult = 0
    bias = [[v + bias_funcs[i](v) for v in new_iso[i]] for i in range(nf)]    # :884
    bcol = [a - b for a, b in zip(bias[c0], bias[c1])]                         # :885
    return SmoothPoints(np.array(midpoints(bcol)), np.array(midpoints(bias[y_index])),
                        np.array(midpoints(color_err)), np.array(midpoints(mag_err[y_index])),
                        np.array(weights), cov_mult, new_mini, new_iso)


def partial_cmd_smooth(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                       completeness_funcs=None, bias_funcs=None, *, dmod=0.0,
                       normalize_value=1.0, mean_mass, ex: Range, ey: Range, nthreads=1,
                       binary_fraction=0.0):
    """partial_cmd_smooth (src/StarFormationHistories.jl:829-911) with
    NoBinaries (binary_fraction 0) or RandomBinaryPairs(binary_fraction).
    Returns (weights matrix (nx, ny), pair count of the single-star part)."""
    sp = partial_cmd_smooth_points(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                                   completeness_funcs, bias_funcs, dmod=dmod,
                                   normalize_value=normalize_value, mean_mass=mean_mass,
                                   ex=ex, ey=ey)
    single, pairs = bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.cov_mult,
                                   sp.weights, ex, ey, nthreads=nthreads)
    if binary_fraction == 0:
        return single, pairs
    ds = 4                                                                     # :898
    binary = binary_hess(sp.new_mini[::ds], [m[::ds] for m in sp.new_iso], mag_err_funcs, y_index,
                         color_indices, imf, completeness_funcs, bias_funcs, ex=ex, ey=ey,
                         normalize_value=normalize_value, mean_mass=mean_mass, nthreads=nthreads)
    b = binary_fraction
    bmf = 2 * b / (1 + b)                                                      # simulate.jl:143,187
    bmf *= trapz(sp.new_mini, [imf(m) for m in sp.new_mini])                   # :907
    return single * (1 - bmf) + bmf * binary, pairs


# --------------------------------------------------------------------------
# Binary systems (src/StarFormationHistories.jl:627-706)
# --------------------------------------------------------------------------

def mag2flux(m, zpt=0.0):
    """mag2flux (src/utilities.jl:57)."""
    return 10.0 ** (4 * (zpt - m) / 10)


def flux2mag(f, zpt=0.0):
    """flux2mag (src/utilities.jl:67)."""
    return zpt - 5 * math.log10(f) / 2


def trapz(x, y):
    """Trapezoidal rule over samples (Trapz.trapz)."""
    return sum((x[i + 1] - x[i]) * (y[i] + y[i + 1]) for i in range(len(x) - 1)) / 2


def _bin_left(v, e: Range):
    """Left-closed bin of v over the edges of e (StatsBase: edges[k] <= v < edges[k+1]), or -1."""
    vals = range_values(e)
    k = bisect.bisect_right(list(vals), v) - 1
    return k if 0 <= k < e.len - 1 else -1


def binary_pairs_hist_loops(m_ini, imf_vec, mag_a, mag_b, y_is_a, prefac, compl, bias, ex: Range, ey: Range):
    """The pair histogram of binary_hess (:645-682), one pair at a time (small
    inputs): returns (nx, ny) weights."""
    nx, ny = ex.len - 1, ey.len - 1
    out = np.zeros((nx, ny))
    dm = [m_ini[i + 1] - m_ini[i] for i in range(len(m_ini) - 1)]
    for i in range(len(dm)):
        for j in range(i + 1, len(dm)):
            ma = flux2mag(mag2flux(mag_a[i]) + mag2flux(mag_a[j]))
            mb = flux2mag(mag2flux(mag_b[i]) + mag2flux(mag_b[j]))
            w = dm[i] * dm[j] * (imf_vec[i] + imf_vec[i + 1]) * (imf_vec[j] + imf_vec[j + 1]) * prefac
            w *= compl[0](ma) * compl[1](mb)
            pa, pb = ma + bias[0](ma), mb + bias[1](mb)
            kx, ky = _bin_left(pa - pb, ex), _bin_left(pa if y_is_a else pb, ey)
            if kx >= 0 and ky >= 0:
                out[kx, ky] += w
    return out


def binary_pairs_hist(m_ini, imf_vec, mag_a, mag_b, y_is_a, prefac, compl, bias, ex: Range, ey: Range):
    """The same, vectorised over the secondary of each primary (numpy, FP64;
    the callables are applied one value at a time)."""
    m_ini = np.asarray(m_ini, dtype=np.float64)
    imf_vec = np.asarray(imf_vec, dtype=np.float64)
    mag_a = np.asarray(mag_a, dtype=np.float64)
    mag_b = np.asarray(mag_b, dtype=np.float64)
    nx, ny = ex.len - 1, ey.len - 1
    ev_x, ev_y = range_values(ex), range_values(ey)
    out = np.zeros(nx * ny)
    n = m_ini.size - 1
    dm = m_ini[1:] - m_ini[:-1]
    isum = imf_vec[:-1] + imf_vec[1:]
    fa, fb = 10.0 ** (4 * (0.0 - mag_a[:n]) / 10), 10.0 ** (4 * (0.0 - mag_b[:n]) / 10)
    ca, cb = np.vectorize(compl[0], otypes=[float]), np.vectorize(compl[1], otypes=[float])
    ba, bb = np.vectorize(bias[0], otypes=[float]), np.vectorize(bias[1], otypes=[float])
    for i in range(n - 1):
        j = np.arange(i + 1, n)
        ma = 0.0 - 5 * np.log10(fa[i] + fa[j]) / 2
        mb = 0.0 - 5 * np.log10(fb[i] + fb[j]) / 2
        w = dm[i] * dm[j] * isum[i] * isum[j] * prefac
        w = w * (ca(ma) * cb(mb))
        pa, pb = ma + ba(ma), mb + bb(mb)
        kx = np.searchsorted(ev_x, pa - pb, side="right") - 1
        ky = np.searchsorted(ev_y, pa if y_is_a else pb, side="right") - 1
        ok = (kx >= 0) & (kx < nx) & (ky >= 0) & (ky < ny)
        np.add.at(out, kx[ok] + nx * ky[ok], w[ok])
    return out.reshape(nx, ny, order="F")


def binary_hess(m_ini, mags, mag_err_funcs, y_index, color_indices, imf, completeness_funcs=None,
                bias_funcs=None, *, ex: Range, ey: Range, normalize_value=1.0, mean_mass, nthreads=1):
    """binary_hess(::RandomBinaryPairs, ...) (:627-706), y_index in
    color_indices: the pair histogram, then bin_cmd_smooth of its non-empty
    bins at their centres.  Returns the (nx, ny) weights."""
    nf = len(mags)
    if completeness_funcs is None:
        completeness_funcs = [lambda m: 1.0] * nf
    if bias_funcs is None:
        bias_funcs = [lambda m: 0.0] * nf
    c0, c1 = color_indices
    assert y_index in color_indices
    imf_vec = [imf(m) for m in m_ini]                                          # :637
    prefac = normalize_value / mean_mass / 4 / trapz(m_ini, imf_vec)           # :642-643
    hist = binary_pairs_hist(m_ini, imf_vec, mags[c0], mags[c1], y_index == c0, prefac,
                             [completeness_funcs[c0], completeness_funcs[c1]],
                             [bias_funcs[c0], bias_funcs[c1]], ex, ey)
    nx, ny = ex.len - 1, ey.len - 1
    hw = hist.ravel(order="F")
    good = np.flatnonzero(hw != 0)                                             # :683
    cx = [ex.first + k * ex.step + ex.step / 2 for k in range(nx)]             # :685-686
    cy = [ey.first + k * ey.step + ey.step / 2 for k in range(ny)]
    h_colors = np.array([cx[g % nx] for g in good])
    h_mags = np.array([cy[g // nx] for g in good])
    ymag_err = np.array([mag_err_funcs[y_index](cy[g // nx]) for g in good])
    xmags = h_mags - h_colors if y_index == c0 else h_mags + h_colors           # :693-697
    x_c_idx = c1 if c0 == y_index else c0
    color_err = np.array([mag_err_funcs[x_c_idx](v) for v in xmags])
    cov_mult = -1 if y_index == c0 else 1
    if good.size == 0:
        return np.zeros((nx, ny))
    w, _ = bin_cmd_smooth(h_colors, h_mags, color_err, ymag_err, cov_mult, hw[good], ex, ey, nthreads=nthreads)
    return w
