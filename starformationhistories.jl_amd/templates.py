"""Smoothed Hess-diagram templates on the GPU, behind the reference's API.

Mirrors the template path of src/StarFormationHistories.jl (v1.3.1):
partial_cmd_smooth (:829-911) prepares the isochrone on the host (numpy,
FP64, the reference's operation order) and hands the kernel centres, widths
and weights to bin_cmd_smooth (:574-621), which runs on the GPU through
libsfh_hip.so.  ``partial_cmd_smooth_batch`` deposits many templates in one
launch and leaves them on the device for the likelihood path.

Index arguments (y_index, color_indices) are 0-based here.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, sfh_range
from .core import Context, HIPTemplates, default_context

# ------------------------------------------------------------------ geometry


@dataclass(frozen=True)
class Range:
    """A uniform range (Julia StepRangeLen): ``first + k*step`` for k < length."""
    first: float
    step: float
    length: int

    def values(self) -> np.ndarray:
        return self.first + np.arange(self.length, dtype=np.float64) * self.step

    def __len__(self):
        return self.length

    @property
    def last(self):
        return self.first + (self.length - 1) * self.step

    def as_c(self) -> sfh_range:
        return sfh_range(float(self.first), float(self.step), int(self.length))


def edges_range(start: float, stop: float, length: int) -> Range:
    """``range(start, stop; length)`` (Float64).  Julia keeps the step in twice
    precision; (stop-start)/(length-1) agrees to ~1 ulp."""
    if length < 2:
        raise ArgumentError("range needs length >= 2 here")
    return Range(float(start), (float(stop) - float(start)) / (length - 1), int(length))


def calculate_edges(edges=None, xlim=None, ylim=None, nbins=None, xwidth=None, ywidth=None):
    """calculate_edges (src/StarFormationHistories.jl:451-476).  ``nbins`` is
    the number of EDGES per axis, as in the reference's range(...; length)."""
    if edges is not None:
        ex, ey = edges
        if not (isinstance(ex, Range) and isinstance(ey, Range)):
            raise ArgumentError("`edges` must be a tuple of two Range objects")
        return edges
    xlim, ylim = sorted(xlim), sorted(ylim)
    if nbins is None:
        if xwidth is not None and ywidth is not None:
            nbins = (int(round((xlim[1] - xlim[0]) / xwidth)), int(round((ylim[1] - ylim[0]) / ywidth)))
        else:
            raise ArgumentError("If the keyword arguments `edges` and `nbins` are not provided, then "
                                "`xwidth` and `ywidth` must be provided.")
    return (edges_range(xlim[0], xlim[1], nbins[0]), edges_range(ylim[0], ylim[1], nbins[1]))


def histogram_pix(d, edges: Range):
    """histogram_pix(d, edges::AbstractRange) (:502): fractional 1-based index."""
    return (np.asarray(d, dtype=np.float64) - edges.first) / edges.step + 1.0


def histogram_data(x, edges: Range):
    """histogram_data(x, edges::AbstractRange) (:518)."""
    return (np.asarray(x, dtype=np.float64) - 1.0) * edges.step + edges.first


@dataclass
class Histogram:
    """The StatsBase.Histogram fields the reference returns: edges, weights
    (nx x ny, colour fastest), closed=:left, isdensity=false."""
    edges: tuple
    weights: np.ndarray
    closed: str = "left"
    isdensity: bool = False


# --------------------------------------------------------- host preparation

def Martin2016_complete(m, A, m50, rho):
    """Martin2016_complete (src/utilities.jl:205)."""
    return A / (1.0 + np.exp((np.asarray(m, dtype=np.float64) - m50) / rho))


def exp_photerr(m, a, b, c, d):
    """exp_photerr (src/utilities.jl:218): a^(b (m - c)) + d."""
    return np.power(a, b * (np.asarray(m, dtype=np.float64) - c)) + d


def _apply(f, x: np.ndarray) -> np.ndarray:
    """Evaluate a user callable elementwise (vectorised when it accepts arrays)."""
    try:
        y = f(x)
        y = np.asarray(y, dtype=np.float64)
        if y.shape == x.shape:
            return y
        if y.ndim == 0:
            return np.full(x.shape, float(y))
    except Exception:
        pass
    return np.array([float(f(float(v))) for v in x], dtype=np.float64)


def mini_spacing(m_ini, colors, mags, dmag, ret_spacing=False):
    """mini_spacing (src/StarFormationHistories.jl:110-148), vectorised: for
    every consecutive pair farther apart than dmag in the CMD insert
    ceil(d/dmag) equal mass steps, then keep first occurrences (unique)."""
    m_ini = np.asarray(m_ini, dtype=np.float64)
    colors = np.asarray(colors, dtype=np.float64)
    mags = np.asarray(mags, dtype=np.float64)
    if not (m_ini.shape == colors.shape == mags.shape):
        raise ArgumentError("axes(m_ini) == axes(mags) == axes(colors)")
    first = m_ini[0]                                  # :117, before sorting
    idx = np.argsort(m_ini, kind="stable")
    m, c, y = m_ini[idx], colors[idx], mags[idx]
    d = np.hypot(np.diff(c), np.diff(y))
    need = d > dmag
    npts = np.where(need, np.ceil(d / dmag), 1.0).astype(np.int64)
    seg = np.repeat(np.arange(d.size), npts)
    j = (np.arange(seg.size) - np.repeat(np.cumsum(npts) - npts, npts) + 1).astype(np.float64)
    mass_step = (m[1:] - m[:-1]) / npts
    vals = np.where(need[seg], m[seg] + mass_step[seg] * j, m[seg + 1])
    new = np.concatenate([[first], vals])
    _, first_idx = np.unique(new, return_index=True)  # :141 unique keeps first occurrences
    new = new[np.sort(first_idx)]
    if not ret_spacing:
        return new
    return new, np.diff(new)


def interpolate_mini(m_ini, mags, new_mini):
    """interpolate_mini (:77-81): sort_ingested then Gridded(Linear())
    interpolation; raises for duplicate knots or out-of-bounds queries
    (the reference's BoundsError)."""
    m_ini = np.asarray(m_ini, dtype=np.float64)
    mags = np.asarray(mags, dtype=np.float64)
    order = np.argsort(m_ini, kind="stable")
    k, yv = m_ini[order], mags[order]
    if np.any(np.diff(k) <= 0):
        raise ArgumentError("duplicate isochrone masses are not supported")
    x = np.asarray(new_mini, dtype=np.float64)
    if x.size and (x.min() < k[0] or x.max() > k[-1]):
        raise ArgumentError("BoundsError: interpolation outside the isochrone's mass range")
    i = np.clip(np.searchsorted(k, x, side="right") - 1, 0, k.size - 2)
    t = (x - k[i]) / (k[i + 1] - k[i])
    return yv[i] * (1.0 - t) + yv[i + 1] * t


_RTOL = math.sqrt(np.finfo(np.float64).eps)


def midpoints(v, ranges: bool = True) -> np.ndarray:
    """midpoints(v, ranges=true) (:417-431); a Range input keeps its step (:436)."""
    if isinstance(v, Range):
        return Range(v.first + v.step / 2, v.step, v.length - 1)
    v = np.asarray(v, dtype=np.float64)
    d = np.diff(v)
    if ranges and d.size and np.all((d == d[0]) | (np.abs(d - d[0]) <= _RTOL * np.maximum(abs(d[0]), np.abs(d)))):
        start, stop, n = v[0] + d[0] / 2, v[-1] - d[0] / 2, v.size - 1
        if n == 1:
            return np.array([start])
        return start + np.arange(n) * ((stop - start) / (n - 1))
    return v[:-1] + d / 2


def calculate_weights(mini, completeness, imf, normalize_value, mean_mass, spacing=None):
    """calculate_weights (:765-776): trapezoid of the IMF pdf across each mass
    step times the completeness of the LEFT point, times normalize/mean_mass."""
    mini = np.asarray(mini, dtype=np.float64)
    completeness = np.asarray(completeness, dtype=np.float64)
    if spacing is None:
        spacing = np.diff(mini)
    if mini.size != completeness.size or np.asarray(spacing).size != mini.size - 1:
        raise ArgumentError("length(mini) == length(completeness) and length(spacing) == length(mini)-1")
    pdf = _apply(imf, mini)
    w = np.asarray(spacing) * (pdf[:-1] + pdf[1:]) / 2
    return w * (completeness[:-1] * normalize_value / mean_mass)


def imf_mean(imf) -> float:
    """mean(imf) (:166-167): the IMF's own mean if it has one."""
    if hasattr(imf, "mean"):
        return float(imf.mean())
    raise ArgumentError("mean_mass must be given for an IMF callable without .mean()")


@dataclass
class SmoothPoints:
    """What partial_cmd_smooth hands to bin_cmd_smooth (:886-888), plus the
    resampled isochrone (new_mini, new_iso_mags) its binary branch uses."""
    colors: np.ndarray
    mags: np.ndarray
    color_err: np.ndarray
    mag_err: np.ndarray
    weights: np.ndarray
    cov_mult: int
    new_mini: np.ndarray = None
    new_iso_mags: list = None


def partial_cmd_smooth_points(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                              completeness_funcs=None, bias_funcs=None, *, dmod=0.0,
                              normalize_value=1.0, mean_mass=None, edges) -> SmoothPoints:
    """Host half of partial_cmd_smooth (:839-888), NoBinaries."""
    nf = len(mags)
    if completeness_funcs is None:
        completeness_funcs = [lambda m: np.ones_like(np.asarray(m, dtype=np.float64))] * nf
    if bias_funcs is None:
        bias_funcs = [lambda m: np.zeros_like(np.asarray(m, dtype=np.float64))] * nf
    if len(color_indices) != 2:
        raise ArgumentError("length(color_indices) == 2")
    if not (nf == len(mag_err_funcs) == len(completeness_funcs) == len(bias_funcs)):
        raise ArgumentError("length(mags) == length(mag_err_funcs) == length(completeness_funcs) == length(bias_funcs)")
    for i in set([y_index, *color_indices]):
        if not (0 <= i < nf):
            raise ArgumentError("y_index / color_indices out of range")
    if mean_mass is None:
        mean_mass = imf_mean(imf)
    ex, ey = edges
    c0, c1 = color_indices
    mags = [np.asarray(m, dtype=np.float64) for m in mags]
    ymags = mags[y_index]
    colors = mags[c0] - mags[c1]                                            # :849
    dmag = min(ex.step, ey.step)                                             # :851
    new_mini, new_spacing = mini_spacing(m_ini, colors, ymags, dmag, True)   # :852
    new_iso = [interpolate_mini(m_ini, m, new_mini) + dmod for m in mags]   # :854
    mag_err = [_apply(mag_err_funcs[i], new_iso[i]) for i in range(nf)]     # :855
    if y_index in color_indices:                                             # :856
        x_c_idx = c1 if c0 == y_index else c0
        color_err = mag_err[x_c_idx]
        compl = _apply(completeness_funcs[c0], new_iso[c0]) * _apply(completeness_funcs[c1], new_iso[c1])
        weights = calculate_weights(new_mini, compl, imf, normalize_value, mean_mass, new_spacing)
        cov_mult = -1 if y_index == c0 else 1                                # :868
    else:
        color_err = np.sqrt(mag_err[c0] ** 2 + mag_err[c1] ** 2)             # :871
        compl = (_apply(completeness_funcs[c0], new_iso[c0]) * _apply(completeness_funcs[c1], new_iso[c1]) *
                 _apply(completeness_funcs[y_index], new_iso[y_index]))
        weights = calculate_weights(new_mini, compl, imf, normalize_value, mean_mass, new_spacing)
        cov_mult = 0
    bias = [new_iso[i] + _apply(bias_funcs[i], new_iso[i]) for i in range(nf)]   # :884
    bcol = bias[c0] - bias[c1]                                               # :885
    return SmoothPoints(midpoints(bcol), midpoints(bias[y_index]), midpoints(color_err),
                        midpoints(mag_err[y_index]), np.asarray(weights, dtype=np.float64), cov_mult,
                        new_mini, new_iso)


# ------------------------------------------------------------- GPU deposition

def bin_cmd_smooth(colors, mags, color_err, mag_err, cov_mult: int = 0, *, weights=None, edges=None,
                   xlim=None, ylim=None, nbins=None, xwidth=None, ywidth=None,
                   ctx: Context | None = None) -> Histogram:
    """bin_cmd_smooth (src/StarFormationHistories.jl:574-621) on the GPU."""
    if cov_mult not in (-1, 0, 1):
        raise ArgumentError("cov_mult in (-1, 0, 1)")
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (colors, mags, color_err, mag_err)]
    if weights is None:
        weights = np.ones_like(arrs[0])
    w = np.ascontiguousarray(weights, dtype=np.float64).ravel()
    if not all(a.size == w.size for a in arrs):
        raise ArgumentError("axes(colors) == axes(mags) == axes(color_err) == axes(mag_err) == axes(weights)")
    if edges is None:
        xlim = xlim if xlim is not None else (arrs[0].min(), arrs[0].max())
        ylim = ylim if ylim is not None else (arrs[1].min(), arrs[1].max())
    ex, ey = calculate_edges(edges, xlim, ylim, nbins, xwidth, ywidth)
    nx, ny = ex.length - 1, ey.length - 1
    out = np.empty(nx * ny, dtype=np.float64)
    ctx = ctx or default_context()
    ctx.check(ctx.lib.sfh_bin_cmd_smooth(ctx.handle, arrs[0].ctypes.data, arrs[1].ctypes.data,
                                         arrs[2].ctypes.data, arrs[3].ctypes.data, w.ctypes.data, w.size,
                                         int(cov_mult), ex.as_c(), ey.as_c(), out.ctypes.data))
    return Histogram((ex, ey), out.reshape((nx, ny), order="F"))


def partial_cmd_smooth(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                       completeness_funcs=None, bias_funcs=None, *, dmod=0.0, normalize_value=1.0,
                       binary_model=None, mean_mass=None, edges=None, xlim=None, ylim=None, nbins=None,
                       xwidth=None, ywidth=None, ctx: Context | None = None) -> Histogram:
    """partial_cmd_smooth (src/StarFormationHistories.jl:829-911): host
    preparation + GPU deposition; with a binary model of non-zero fraction the
    single-star template is mixed with binary_hess of the isochrone
    downsampled 4x (:895-910)."""
    binary_model = NoBinaries() if binary_model is None else binary_model
    edges = calculate_edges(edges, xlim, ylim, nbins, xwidth, ywidth)
    if mean_mass is None:
        mean_mass = imf_mean(imf)
    sp = partial_cmd_smooth_points(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
                                   completeness_funcs, bias_funcs, dmod=dmod,
                                   normalize_value=normalize_value, mean_mass=mean_mass, edges=edges)
    single = bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.cov_mult, weights=sp.weights,
                            edges=edges, ctx=ctx)
    bfrac = binary_system_fraction(binary_model)
    if not bfrac <= 1:
        raise ArgumentError("binary_system_fraction(binary_model) <= 1")
    if bfrac == 0:
        return single
    ds = 4                                                                   # :898
    binary = binary_hess(binary_model, sp.new_mini[::ds], [m[::ds] for m in sp.new_iso_mags], mag_err_funcs,
                         y_index, color_indices, imf, completeness_funcs, bias_funcs, edges,
                         normalize_value=normalize_value, mean_mass=mean_mass, ctx=ctx)
    bmf = binary_mass_fraction(binary_model, imf)
    bmf *= trapz(sp.new_mini, _apply(imf, sp.new_mini))                      # :907
    return Histogram(edges, single.weights * (1 - bmf) + bmf * binary.weights)


# ------------------------------------------------------------------ binaries

class NoBinaries:
    """NoBinaries (src/simulate.jl:150-157)."""
    fraction = 0.0


class RandomBinaryPairs:
    """RandomBinaryPairs(fraction) (src/simulate.jl:159-170): binaries as
    random pairs drawn from the single-star IMF, 0 <= fraction <= 1."""

    def __init__(self, fraction: float):
        if not (0.0 <= fraction <= 1.0):
            raise ArgumentError("(fraction >= 0) && (fraction <= 1)")
        self.fraction = float(fraction)


def binary_system_fraction(model) -> float:
    """src/simulate.jl:139: the model's fraction (0 for NoBinaries)."""
    return float(model.fraction)


def binary_number_fraction(model) -> float:
    """src/simulate.jl:143: 2b / (1 + b)."""
    b = binary_system_fraction(model)
    return 2 * b / (1 + b)


def binary_mass_fraction(model, imf) -> float:
    """src/simulate.jl:157,187: 0 for NoBinaries, the number fraction of
    stars in pairs for RandomBinaryPairs."""
    if isinstance(model, NoBinaries):
        return 0.0
    if isinstance(model, RandomBinaryPairs):
        return binary_number_fraction(model)
    raise NotImplementedError(f"binary_mass_fraction for {type(model).__name__}")


def mag2flux(m, zpt=0.0):
    """mag2flux (src/utilities.jl:57): exp10(4 (zpt - m) / 10)."""
    return np.power(10.0, 4.0 * (zpt - np.asarray(m, dtype=np.float64)) / 10.0)


def flux2mag(f, zpt=0.0):
    """flux2mag (src/utilities.jl:67): zpt - 5 log10(f) / 2."""
    return zpt - 5.0 * np.log10(np.asarray(f, dtype=np.float64)) / 2.0


def trapz(x, y) -> float:
    """Trapezoidal integral of samples y(x) (Trapz.trapz)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return float(np.sum((x[1:] - x[:-1]) * (y[1:] + y[:-1])) / 2)


_TABLE_POINTS = 16385


def _table(f, lo, hi):
    """f tabulated on a uniform grid over [lo, hi] (an sfh_table) and the
    array that backs it."""
    grid = np.linspace(lo, hi, _TABLE_POINTS)
    vals = np.ascontiguousarray(_apply(f, grid))
    return _lib.sfh_table(float(grid[0]), float(grid[1] - grid[0]), vals.size, vals.ctypes.data), vals


def binary_pairs_hist(m_ini, imf_pdf, mag_a, mag_b, y_is_a, prefac, completeness_funcs, bias_funcs, edges,
                      ctx: Context | None = None) -> np.ndarray:
    """The pair histogram of binary_hess (:627-682) on the GPU
    (sfh_binary_pairs_hist): (nx, ny) FP64.  completeness_funcs / bias_funcs:
    the callables of colour filters a and b, tabulated here (16385 points over
    the pairs' magnitude range, linear interpolation)."""
    ctx = ctx or default_context()
    m_ini = np.ascontiguousarray(m_ini, dtype=np.float64)
    imf_pdf = np.ascontiguousarray(imf_pdf, dtype=np.float64)
    mag_a = np.ascontiguousarray(mag_a, dtype=np.float64)
    mag_b = np.ascontiguousarray(mag_b, dtype=np.float64)
    n = m_ini.size
    if not (imf_pdf.size == mag_a.size == mag_b.size == n):
        raise ArgumentError("all(length(i) == length(m_ini) for i in mags)")
    ex, ey = edges
    out = np.zeros((ex.length - 1) * (ey.length - 1))
    if n < 3:
        return out.reshape(ex.length - 1, ey.length - 1, order="F")
    tabs, keep = [], []
    for mags, f_c, f_b in ((mag_a, completeness_funcs[0], bias_funcs[0]), (mag_b, completeness_funcs[1], bias_funcs[1])):
        # a pair is at most 2.5 log10(2) brighter than its brighter member
        lo, hi = float(mags[:-1].min()) - 0.76, float(mags[:-1].max()) + 1e-9
        for f in (f_c, f_b):
            t, v = _table(f, lo, hi)
            tabs.append(t)
            keep.append(v)
    ctx.check(ctx.lib.sfh_binary_pairs_hist(ctx.handle, n, m_ini.ctypes.data, imf_pdf.ctypes.data,
                                            mag_a.ctypes.data, mag_b.ctypes.data, 1 if y_is_a else 0,
                                            float(prefac), C.byref(tabs[0]), C.byref(tabs[2]), C.byref(tabs[1]),
                                            C.byref(tabs[3]), ex.as_c(), ey.as_c(), out.ctypes.data))
    return out.reshape(ex.length - 1, ey.length - 1, order="F")


def binary_hess(model, m_ini, mags, mag_err_funcs, y_index, color_indices, imf, completeness_funcs=None,
                bias_funcs=None, edges=None, *, normalize_value=1.0, mean_mass=None,
                ctx: Context | None = None) -> Histogram:
    """binary_hess(::RandomBinaryPairs, ...) (src/StarFormationHistories.jl:627-706):
    the pair histogram on the GPU, then bin_cmd_smooth of its non-empty bins
    at their centres (GPU)."""
    if not isinstance(model, RandomBinaryPairs):
        raise NotImplementedError(f"binary_hess(model::{type(model).__name__}, ...) not implemented.")
    nf = len(mags)
    if completeness_funcs is None:
        completeness_funcs = [lambda m: np.ones_like(np.asarray(m, dtype=np.float64))] * nf
    if bias_funcs is None:
        bias_funcs = [lambda m: np.zeros_like(np.asarray(m, dtype=np.float64))] * nf
    m_ini = np.asarray(m_ini, dtype=np.float64)
    mags = [np.asarray(m, dtype=np.float64) for m in mags]
    if not all(m.size == m_ini.size for m in mags):
        raise ArgumentError("all(length(i) == length(m_ini) for i in mags)")
    if mean_mass is None:
        mean_mass = imf_mean(imf)
    c0, c1 = color_indices
    if y_index not in color_indices:                                         # :699-701
        raise NotImplementedError("`binary_hess(model::RandomBinaryPairs, ...)` for "
                                  "`!(y_index in color_indices)` not yet implemented.")
    ex, ey = edges
    imf_vec = _apply(imf, m_ini)                                             # :637
    prefac = normalize_value / mean_mass / 4 / trapz(m_ini, imf_vec)          # :642-643
    hist = binary_pairs_hist(m_ini, imf_vec, mags[c0], mags[c1], y_index == c0, prefac,
                             [completeness_funcs[c0], completeness_funcs[c1]], [bias_funcs[c0], bias_funcs[c1]],
                             edges, ctx=ctx)
    hw = hist.ravel(order="F")                                               # vec(hist.weights)
    good = np.flatnonzero(hw != 0)                                           # :683
    nx, ny = ex.length - 1, ey.length - 1
    cx = ex.first + np.arange(nx) * ex.step + ex.step / 2                    # :685-686
    cy = ey.first + np.arange(ny) * ey.step + ey.step / 2
    h_colors = np.tile(cx, ny)[good]
    h_mags = np.repeat(cy, nx)[good]
    ymag_err = np.repeat(_apply(mag_err_funcs[y_index], cy), nx)[good]
    xmags = h_mags - h_colors if y_index == c0 else h_mags + h_colors          # :693-697
    x_c_idx = c1 if c0 == y_index else c0
    color_err = _apply(mag_err_funcs[x_c_idx], xmags)
    cov_mult = -1 if y_index == c0 else 1
    if good.size == 0:
        return Histogram(edges, np.zeros((nx, ny)))
    return bin_cmd_smooth(h_colors, h_mags, color_err, ymag_err, cov_mult, weights=hw[good], edges=edges, ctx=ctx)


def binary_pairs_hist_batch(isochrones, imf, mag_filters, y_is_a, prefacs, completeness_funcs, bias_funcs, edges,
                            ctx: Context | None = None) -> np.ndarray:
    """The pair histograms of binary_hess for many isochrones in one launch
    (sfh_binary_pairs_hist_batch): ``isochrones`` is a list of (m_ini, mags)
    with mags indexed by filter, ``mag_filters`` the colour filters (a, b).
    The completeness / bias callables are tabulated once over all the pairs'
    magnitudes.  Returns (ntemplates, nx, ny) FP64."""
    ctx = ctx or default_context()
    a, b = mag_filters
    ms = [np.asarray(m, dtype=np.float64) for m, _ in isochrones]
    ga = [np.asarray(g[a], dtype=np.float64) for _, g in isochrones]
    gb = [np.asarray(g[b], dtype=np.float64) for _, g in isochrones]
    if not all(x.size == y.size == z.size for x, y, z in zip(ms, ga, gb)):
        raise ArgumentError("all(length(i) == length(m_ini) for i in mags)")
    nt = len(ms)
    ex, ey = edges
    nx, ny = ex.length - 1, ey.length - 1
    out = np.zeros(nt * nx * ny)
    if nt == 0:
        return out.reshape(0, nx, ny)
    imfv = [np.asarray(_apply(imf, m), dtype=np.float64) for m in ms]
    offsets = np.concatenate([[0], np.cumsum([m.size for m in ms])]).astype(np.int64)
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs), dtype=np.float64)
    m_all, i_all, a_all, b_all = cat(ms), cat(imfv), cat(ga), cat(gb)
    pre = np.ascontiguousarray(prefacs, dtype=np.float64)
    tabs, keep = [], []
    for mags, f_c, f_b in ((ga, completeness_funcs[0], bias_funcs[0]), (gb, completeness_funcs[1], bias_funcs[1])):
        body = [g[:-1] for g in mags if g.size > 1]
        if body:
            lo, hi = min(float(g.min()) for g in body) - 0.76, max(float(g.max()) for g in body) + 1e-9
        else:
            lo, hi = 0.0, 1.0
        for f in (f_c, f_b):
            t, v = _table(f, lo, hi)
            tabs.append(t)
            keep.append(v)
    ctx.check(ctx.lib.sfh_binary_pairs_hist_batch(ctx.handle, nt, offsets.ctypes.data, m_all.ctypes.data,
                                                  i_all.ctypes.data, a_all.ctypes.data, b_all.ctypes.data,
                                                  1 if y_is_a else 0, pre.ctypes.data, C.byref(tabs[0]),
                                                  C.byref(tabs[2]), C.byref(tabs[1]), C.byref(tabs[3]), ex.as_c(),
                                                  ey.as_c(), out.ctypes.data))
    return out.reshape(nt, ny, nx).transpose(0, 2, 1)


class _BinPoints:
    """Kernel points of a binary template (bin centres of its pair histogram)."""

    def __init__(self, colors, mags, color_err, mag_err, weights, cov_mult):
        self.colors, self.mags, self.color_err, self.mag_err = colors, mags, color_err, mag_err
        self.weights, self.cov_mult = weights, cov_mult


def binary_hess_batch(model, isochrones, mag_err_funcs, y_index, color_indices, imf, completeness_funcs=None,
                      bias_funcs=None, edges=None, *, normalize_value=1.0, mean_mass=None, out=None,
                      ctx: Context | None = None):
    """binary_hess(::RandomBinaryPairs, ...) (src/StarFormationHistories.jl:627-706)
    for many isochrones ((m_ini, mags) pairs, mags indexed by filter): every
    pair histogram in one launch, then the non-empty bins of all of them
    deposited in one batched launch.  Returns an FP32 device tensor
    (ntemplates x nbins), as deposit_batch does."""
    import torch
    if not isinstance(model, RandomBinaryPairs):
        raise NotImplementedError(f"binary_hess(model::{type(model).__name__}, ...) not implemented.")
    ctx = ctx or default_context()
    c0, c1 = color_indices
    if y_index not in color_indices:
        raise NotImplementedError("`binary_hess(model::RandomBinaryPairs, ...)` for "
                                  "`!(y_index in color_indices)` not yet implemented.")
    nf = len(isochrones[0][1]) if isochrones else 2
    if completeness_funcs is None:
        completeness_funcs = [lambda m: np.ones_like(np.asarray(m, dtype=np.float64))] * nf
    if bias_funcs is None:
        bias_funcs = [lambda m: np.zeros_like(np.asarray(m, dtype=np.float64))] * nf
    if mean_mass is None:
        mean_mass = imf_mean(imf)
    ex, ey = edges
    nx, ny = ex.length - 1, ey.length - 1
    pre = [normalize_value / mean_mass / 4 / trapz(m, _apply(imf, np.asarray(m, dtype=np.float64)))
           for m, _ in isochrones]                                           # :642-643
    hist = binary_pairs_hist_batch(isochrones, imf, (c0, c1), y_index == c0, pre,
                                   [completeness_funcs[c0], completeness_funcs[c1]],
                                   [bias_funcs[c0], bias_funcs[c1]], edges, ctx=ctx)
    cx = ex.first + np.arange(nx) * ex.step + ex.step / 2                    # :685-686
    cy = ey.first + np.arange(ny) * ey.step + ey.step / 2
    x_c_idx = c1 if c0 == y_index else c0
    cov_mult = -1 if y_index == c0 else 1
    ycol = _apply(mag_err_funcs[y_index], cy)
    pts = []
    for t in range(len(isochrones)):
        hw = hist[t].ravel(order="F")
        good = np.flatnonzero(hw != 0)                                       # :683
        h_colors = np.tile(cx, ny)[good]
        h_mags = np.repeat(cy, nx)[good]
        xmags = h_mags - h_colors if y_index == c0 else h_mags + h_colors    # :693-697
        color_err = _apply(mag_err_funcs[x_c_idx], xmags) if good.size else np.zeros(0)
        pts.append(_BinPoints(h_colors, h_mags, color_err, np.repeat(ycol, nx)[good], hw[good], cov_mult))
    nt = len(pts)
    if out is None:
        out = torch.zeros((nt, nx * ny), dtype=torch.float32, device="cuda")
    if sum(p.colors.size for p in pts) == 0:
        out.zero_()
        torch.cuda.synchronize()
        return out
    deposit_batch(PointBatch(pts, edges), out=out, ctx=ctx)
    ctx.synchronize()
    return out


class PointBatch:
    """Concatenated kernel points of many templates, resident on the device
    (torch tensors), ready for ``deposit_batch``."""

    def __init__(self, points_list, edges, cov_mult=None, cmult_list=None, device="cuda"):
        import torch
        self.edges = edges
        self.ntemplates = len(points_list)
        counts = [p.colors.size for p in points_list]
        self.offsets_host = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        cat = lambda attr: np.concatenate([getattr(p, attr) for p in points_list]).astype(np.float64)
        self.cov_mult = float(points_list[0].cov_mult if cov_mult is None else cov_mult)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.colors, self.mags = t(cat("colors")), t(cat("mags"))
        self.color_err, self.mag_err, self.weights = t(cat("color_err")), t(cat("mag_err")), t(cat("weights"))
        self.cmult = t(np.concatenate(cmult_list).astype(np.float64)) if cmult_list is not None else None
        self.offsets = t(self.offsets_host)

    @property
    def npoints(self):
        return int(self.offsets_host[-1])


def deposit_batch(batch: PointBatch, out=None, ctx: Context | None = None, return_pairs=False):
    """Deposit every template of ``batch`` in one launch
    (sfh_bin_cmd_smooth_batch_device).  Returns an FP32 device tensor
    (nbins x ntemplates column-major, i.e. torch shape (ntemplates, nbins))."""
    import torch
    ctx = ctx or default_context()
    ex, ey = batch.edges
    nbins = (ex.length - 1) * (ey.length - 1)
    if out is None:
        out = torch.empty((batch.ntemplates, nbins), dtype=torch.float32, device=batch.colors.device)
    pairs = C.c_int64()
    ctx.check(ctx.lib.sfh_bin_cmd_smooth_batch_device(
        ctx.handle, batch.colors.data_ptr(), batch.mags.data_ptr(), batch.color_err.data_ptr(),
        batch.mag_err.data_ptr(), batch.weights.data_ptr(),
        batch.cmult.data_ptr() if batch.cmult is not None else None, batch.offsets.data_ptr(),
        batch.ntemplates, batch.cov_mult, ex.as_c(), ey.as_c(), out.data_ptr(), nbins,
        C.byref(pairs) if return_pairs else None))
    if return_pairs:
        return out, pairs.value
    return out


def templates_from_device(out, nbins: int, ctx: Context | None = None) -> HIPTemplates:
    """Wrap deposited device templates as the likelihood path's HIPTemplates."""
    return HIPTemplates(ctx=ctx, device_ptr=out.data_ptr(), shape=(nbins, out.shape[0]))


# ------------------------------------------ partial_cmd_smooth prepared on the device

class ExpPhotErr:
    """m -> exp_photerr(m, a, b, c, d) (src/utilities.jl:218) as a callable
    the device evaluates exactly."""

    def __init__(self, a, b, c, d):
        self.a, self.b, self.c, self.d = float(a), float(b), float(c), float(d)

    def __call__(self, m):
        return exp_photerr(m, self.a, self.b, self.c, self.d)

    def as_fn(self, lo=None, hi=None):
        return _lib.sfh_fn(_lib.SFH_FN_EXP_PHOTERR, (C.c_double * 4)(self.a, self.b, self.c, self.d))


class Martin2016Completeness:
    """m -> Martin2016_complete(m, A, m50, rho) (src/utilities.jl:205) as a
    callable the device evaluates exactly."""

    def __init__(self, A, m50, rho):
        self.A, self.m50, self.rho = float(A), float(m50), float(rho)

    def __call__(self, m):
        return Martin2016_complete(m, self.A, self.m50, self.rho)

    def as_fn(self, lo=None, hi=None):
        return _lib.sfh_fn(_lib.SFH_FN_MARTIN2016, (C.c_double * 4)(self.A, self.m50, self.rho, 0.0))


class Constant:
    """m -> value (the reference's default `one` / `zero` callables)."""

    def __init__(self, value):
        self.value = float(value)

    def __call__(self, m):
        return np.full(np.shape(m), self.value)

    def as_fn(self, lo=None, hi=None):
        return _lib.sfh_fn(_lib.SFH_FN_CONST, (C.c_double * 4)(self.value, 0.0, 0.0, 0.0))


class Tabulated:
    """Any callable f, tabulated on n points over [lo, hi] and interpolated
    linearly by the device (clamped outside): the route for user callables the
    device cannot evaluate directly.  The interpolation error is at most
    h^2/8 max|f''| with h = (hi - lo)/(n - 1)."""

    def __init__(self, f, lo, hi, n=16385):
        if not (hi > lo and n >= 2):
            raise ArgumentError("Tabulated needs hi > lo and n >= 2")
        self.f, self.lo, self.hi, self.n = f, float(lo), float(hi), int(n)
        self.step = (self.hi - self.lo) / (self.n - 1)
        self.values = np.ascontiguousarray(_apply(f, self.lo + self.step * np.arange(self.n)), dtype=np.float64)

    def __call__(self, m):
        return np.interp(np.asarray(m, dtype=np.float64), self.lo + self.step * np.arange(self.n), self.values)

    def as_fn(self, lo=None, hi=None):
        fn = _lib.sfh_fn(_lib.SFH_FN_TABLE)
        fn.table = _lib.sfh_table(self.lo, self.step, self.n, self.values.ctypes.data)
        return fn


def _fn_of(f, arg_range, keep):
    """The sfh_fn of callable ``f``: its own form when it has one
    (``as_fn``), else a table over the arguments it will see, [lo, hi] =
    ``arg_range()`` (called only then), kept alive in ``keep``."""
    if hasattr(f, "as_fn"):
        keep.append(f)
        return f.as_fn()
    lo, hi = arg_range()
    if not hi > lo:
        hi = lo + 1.0
    # exactly over [lo, hi]: arguments are clamped to the table, and a padded
    # range would sample functions with hard limits (an IMF's mass range) outside
    # their support
    t = Tabulated(f, lo, hi)
    keep.append(t)
    return t.as_fn()


def _pinned_f64(n):
    """A float64 host array of n elements in page-locked memory when a GPU is
    present (the library's copies from it are direct DMA at the full bus
    rate, where pageable memory goes through a staging copy), else an
    ordinary array.  Returns (array, owner): keep the owner alive with it."""
    try:
        import torch
        if torch.cuda.is_available():
            t = torch.empty(max(int(n), 1), dtype=torch.float64, pin_memory=True)
            return t.numpy()[:n], t
    except (ImportError, RuntimeError):
        pass
    return np.empty(n, dtype=np.float64), None


class IsochroneBatch:
    """Many isochrones in the packed form sfh_partial_cmd_smooth_batch takes:
    isochrone t owns entries offsets[t] .. offsets[t+1]-1 of m_ini and of each
    filter's magnitude array.  Building it once lets repeated batch calls skip
    the per-call concatenation of per-isochrone arrays; the packed arrays live
    in page-locked host memory when a GPU is present, so each call's upload of
    them is a direct DMA."""

    def __init__(self, m_ini, mags, offsets):
        self._owners = []
        self.m_ini = self._pack(np.asarray(m_ini, dtype=np.float64).ravel())
        self.mags = [self._pack(np.asarray(a, dtype=np.float64).ravel()) for a in mags]
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64).ravel()
        n = self.m_ini.size
        if self.offsets.size < 2 or self.offsets[0] != 0 or self.offsets[-1] != n or np.any(np.diff(self.offsets) < 0):
            raise ArgumentError("offsets must run from 0 to length(m_ini), non-decreasing")
        if any(a.size != n for a in self.mags):
            raise ArgumentError("axes(m_ini) == axes(mags[i])")

    def _pack(self, a):
        out, owner = _pinned_f64(a.size)
        out[:] = a
        if owner is not None:
            self._owners.append(owner)
        return out

    @classmethod
    def from_list(cls, isochrones):
        """From a sequence of (m_ini, mags) pairs."""
        m_cat = [np.asarray(m, dtype=np.float64).ravel() for m, _ in isochrones]
        counts = np.array([m.size for m in m_cat], dtype=np.int64)
        nf = len(isochrones[0][1]) if len(isochrones) else 0
        arrs = [[np.asarray(mg[f], dtype=np.float64).ravel() for _, mg in isochrones] for f in range(nf)]
        for per_filter in arrs:
            if any(a.size != n for a, n in zip(per_filter, counts)):
                raise ArgumentError("axes(m_ini) == axes(mags[i])")
        # concatenated straight into the packed (page-locked) arrays
        self = cls.__new__(cls)
        self._owners = []
        n = int(counts.sum())

        def packed(parts):
            out, owner = _pinned_f64(n)
            if owner is not None:
                self._owners.append(owner)
            if parts:
                np.concatenate(parts, out=out)
            return out
        self.m_ini = packed(m_cat)
        self.mags = [packed(per_filter) for per_filter in arrs]
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        return self

    def __len__(self):
        return self.offsets.size - 1

    def ranges(self):
        """(min, max) of m_ini and of each filter's magnitudes, computed once
        (for tabulating callables over the arguments they will see)."""
        if getattr(self, "_ranges", None) is None:
            mm = lambda a: (float(a.min()), float(a.max())) if a.size else (0.0, 1.0)  # noqa: E731
            self._ranges = (mm(self.m_ini), [mm(a) for a in self.mags])
        return self._ranges

    @property
    def nfilters(self):
        return len(self.mags)


def partial_cmd_smooth_batch(isochrones, mag_err_funcs, y_index, color_indices, imf, completeness_funcs=None,
                             bias_funcs=None, *, dmod=0.0, normalize_value=1.0, mean_mass=None, correlation=0.0,
                             edges=None, xlim=None, ylim=None, nbins=None, xwidth=None, ywidth=None,
                             output="device", return_info=False, ctx: Context | None = None):
    """partial_cmd_smooth (src/StarFormationHistories.jl:829-893, NoBinaries)
    for many isochrones in one call (sfh_partial_cmd_smooth_batch): the host
    preparation of every isochrone runs on the device, then one deposition
    launch.  ``isochrones``: a sequence of (m_ini, mags) with m_ini strictly
    increasing and mags a sequence of per-filter magnitude arrays, or an
    IsochroneBatch of the same in packed form.  The
    callables are shared by all isochrones; ``dmod``, ``normalize_value`` and
    ``mean_mass`` may be scalars or one value per isochrone.  Callables with an
    ``as_fn`` form (ExpPhotErr, Martin2016Completeness, Constant, Tabulated,
    power-law IMFs) are evaluated exactly; any other callable is tabulated on
    16385 points over the magnitudes (or masses) it will see.
    ``correlation`` r != 0 replaces cov_mult by c = -r/sqrt(1-r^2) color_err/mag_err
    per kernel point (a colour-magnitude correlation, not in the reference).
    Returns a device tensor (ntemplates, nbins) (``output="device"``) or a host
    array nbins x ntemplates (``"host"``); with ``return_info`` also
    (kernel points, pairs)."""
    import torch
    ctx = ctx or default_context()
    batch = isochrones if isinstance(isochrones, IsochroneBatch) else IsochroneBatch.from_list(isochrones)
    T = len(batch)
    if T == 0:
        raise ArgumentError("need at least one isochrone")
    if len(color_indices) != 2:
        raise ArgumentError("length(color_indices) == 2")
    nf = batch.nfilters
    if completeness_funcs is None:
        completeness_funcs = [None] * nf
    if bias_funcs is None:
        bias_funcs = [None] * nf
    if not (nf == len(mag_err_funcs) == len(completeness_funcs) == len(bias_funcs)):
        raise ArgumentError("length(mags) == length(mag_err_funcs) == length(completeness_funcs) == "
                            "length(bias_funcs)")
    for i in (y_index, *color_indices):
        if not 0 <= i < nf:
            raise ArgumentError("y_index / color_indices out of range")
    ex, ey = calculate_edges(edges, xlim, ylim, nbins, xwidth, ywidth)
    offsets, m_all, mags_all = batch.offsets, batch.m_ini, batch.mags
    per = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (T,)))  # noqa: E731
    dm = per(dmod)
    if mean_mass is None:
        mean_mass = imf_mean(imf)
    nv, mm = per(normalize_value), per(mean_mass)
    keep = [m_all, mags_all, dm, nv, mm, offsets]
    # argument ranges the callables will see (the filters' magnitudes plus
    # dmod; the masses), needed only for callables that are tabulated
    mag_range = lambda f: (batch.ranges()[1][f][0] + float(dm.min()), batch.ranges()[1][f][1] + float(dm.max()))  # noqa: E731,E501
    FnArr = _lib.sfh_fn * nf
    err = FnArr(*[_fn_of(mag_err_funcs[f], lambda f=f: mag_range(f), keep) for f in range(nf)])
    comp = FnArr(*[(_fn_of(completeness_funcs[f], lambda f=f: mag_range(f), keep) if completeness_funcs[f] is not None
                    else Constant(1.0).as_fn()) for f in range(nf)])
    bias = FnArr(*[(_fn_of(bias_funcs[f], lambda f=f: mag_range(f), keep) if bias_funcs[f] is not None
                    else Constant(0.0).as_fn()) for f in range(nf)])
    imf_fn = _fn_of(imf, lambda: batch.ranges()[0], keep)
    mag_ptrs = (C.c_void_p * nf)(*[a.ctypes.data for a in mags_all])
    nb = (ex.length - 1) * (ey.length - 1)
    npts, pairs = C.c_int64(), C.c_int64()
    if output == "device":
        out = torch.empty((T, nb), dtype=torch.float32, device="cuda")
        # the allocator may hand back memory that work still queued on torch's
        # stream reads; the library writes it on its own stream
        torch.cuda.current_stream().synchronize()
        hptr, dptr = None, out.data_ptr()
    elif output == "host":
        out = np.empty((nb, T), dtype=np.float32, order="F")
        hptr, dptr = out.ctypes.data, None
    else:
        raise ArgumentError('output must be "device" or "host"')
    ctx.check(ctx.lib.sfh_partial_cmd_smooth_batch(
        ctx.handle, T, offsets.ctypes.data, m_all.ctypes.data, nf, mag_ptrs, int(y_index), int(color_indices[0]),
        int(color_indices[1]), err, comp, bias, C.byref(imf_fn), dm.ctypes.data, nv.ctypes.data, mm.ctypes.data,
        float(correlation), ex.as_c(), ey.as_c(), hptr, dptr, nb, C.byref(npts), C.byref(pairs)))
    del keep
    if return_info:
        return out, (npts.value, pairs.value)
    return out
