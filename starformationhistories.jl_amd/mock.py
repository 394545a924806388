"""Mock observations on the GPU: stellar populations sampled from isochrones
(generate_stars_mass / generate_stars_mag and their composite forms,
src/simulate.jl:306-499), observed with photometric errors, bias and
completeness (model_cmd, :621-674), and binned into Hess diagrams (bin_cmd,
src/StarFormationHistories.jl:544-553).  Host preparation (ingest_mags,
sort_ingested, mass_limits) follows the reference; sampling, observing and
binning run in libsfh_hip.so (sfh_generate_stars, sfh_model_cmd,
sfh_bin_cmd).

Random numbers are counter based (Philox4x64-10 keyed by ``seed`` and
``stream``; system or star i uses counter i), so a result depends only on
(seed, stream) -- the reference's ``rng`` keyword becomes ``seed``.  The IMF
is a piecewise power law (``PowerLawIMF``, ``Salpeter1955``,
``Kroupa2001``); user callables for errors, completeness and bias are
tabulated on 16385 points over the magnitudes they are applied to.
"""
from __future__ import annotations

import ctypes as C
import math
import struct

import numpy as np

from . import _lib
from ._lib import ArgumentError, DomainError
from .core import Context, default_context
from .templates import (Histogram, NoBinaries, RandomBinaryPairs, Range, _apply, calculate_edges, flux2mag,
                        mag2flux, trapz)

_TABLE_POINTS = 16385


# ------------------------------------------------------------------ IMFs

class PowerLawIMF:
    """pdf(m) proportional to coefs[k] m^-slopes[k] on [breaks[k], breaks[k+1])."""

    def __init__(self, breaks, slopes, coefs=None):
        self.breaks = np.ascontiguousarray(breaks, dtype=np.float64)
        self.slopes = np.ascontiguousarray(slopes, dtype=np.float64)
        self.coefs = np.ascontiguousarray(np.ones(self.slopes.size) if coefs is None else coefs, dtype=np.float64)
        if not (self.breaks.size == self.slopes.size + 1 == self.coefs.size + 1 and self.slopes.size >= 1):
            raise ArgumentError("need nseg + 1 breaks, nseg slopes and nseg coefficients")
        if self.slopes.size > 8:
            raise ArgumentError("at most 8 power-law segments")
        if not (np.all(np.diff(self.breaks) > 0) and self.breaks[0] > 0 and np.all(self.coefs > 0)):
            raise ArgumentError("breaks must increase from a positive mass; coefficients positive")
        self._norm = 1.0 / sum(self._int(k, 1.0, *self.breaks[k:k + 2]) for k in range(self.slopes.size))

    def _int(self, k, power, lo, hi):
        """integral of coef_k m^(power - 1 - slope_k) over [lo, hi]"""
        a = power - self.slopes[k]
        if a == 0:
            return self.coefs[k] * math.log(hi / lo)
        return self.coefs[k] * (hi ** a - lo ** a) / a

    def minimum(self):
        return float(self.breaks[0])

    def maximum(self):
        return float(self.breaks[-1])

    def pdf(self, m):
        m = np.asarray(m, dtype=np.float64)
        out = np.zeros(m.shape)
        for k in range(self.slopes.size):
            lo, hi = self.breaks[k], self.breaks[k + 1]
            sel = (m >= lo) & ((m < hi) | ((k == self.slopes.size - 1) & (m == hi)))
            out = np.where(sel, self.coefs[k] * np.power(np.where(sel, m, 1.0), -self.slopes[k]) * self._norm, out)
        return out

    def __call__(self, m):
        return self.pdf(m)

    def cdf(self, m):
        m = np.atleast_1d(np.asarray(m, dtype=np.float64))
        out = np.zeros(m.shape)
        for i, v in enumerate(m):
            s = 0.0
            for k in range(self.slopes.size):
                lo, hi = self.breaks[k], self.breaks[k + 1]
                if v > lo:
                    s += self._int(k, 1.0, lo, min(v, hi))
            out[i] = s * self._norm
        return out

    def mean(self):
        return self._norm * sum(self._int(k, 2.0, *self.breaks[k:k + 2]) for k in range(self.slopes.size))

    def as_c(self):
        return _lib.sfh_imf(self.slopes.size, self.breaks.ctypes.data, self.slopes.ctypes.data,
                            self.coefs.ctypes.data)


def Salpeter1955(mmin, mmax):
    """dN/dm proportional to m^-2.35 on [mmin, mmax]."""
    return PowerLawIMF([mmin, mmax], [2.35])


def Kroupa2001(mmin, mmax):
    """Kroupa (2001) for mmin >= 0.08: m^-1.3 below 0.5 Msun, 0.5 m^-2.3 above."""
    if mmin < 0.08:
        raise ArgumentError("mmin >= 0.08 only")
    if mmax <= 0.5:
        return PowerLawIMF([mmin, mmax], [1.3], [1.0])
    if mmin >= 0.5:
        return PowerLawIMF([mmin, mmax], [2.3], [0.5])
    return PowerLawIMF([mmin, 0.5, mmax], [1.3, 2.3], [1.0, 0.5])


class BinaryMassRatio:
    """BinaryMassRatio(fraction, Uniform(qmin, qmax)) (src/simulate.jl:195-207):
    a system mass from the IMF, split by a mass ratio q = M_s / M_p."""

    def __init__(self, fraction: float, qmin: float = 0.1, qmax: float = 1.0):
        if not (0.0 <= fraction <= 1.0):
            raise ArgumentError("(fraction >= 0) && (fraction <= 1)")
        if not (0.0 <= qmin <= qmax <= 1.0):
            raise ArgumentError("(minimum(qdist) >= 0) && (maximum(qdist) <= 1)")
        self.fraction, self.qmin, self.qmax = float(fraction), float(qmin), float(qmax)


def _binary_c(model):
    if isinstance(model, NoBinaries):
        return _lib.sfh_binary_model(0, 0.0, 0.0, 0.0), 1
    if isinstance(model, RandomBinaryPairs):
        return _lib.sfh_binary_model(1, model.fraction, 0.0, 0.0), 2
    if isinstance(model, BinaryMassRatio):
        return _lib.sfh_binary_model(2, model.fraction, model.qmin, model.qmax), 2
    raise ArgumentError(f"unsupported binary model {type(model).__name__}")


# ------------------------------------------------------------------ ingestion

def ingest_mags(mini_vec, mags) -> np.ndarray:
    """ingest_mags (src/simulate.jl:12-47): magnitudes as an (nstars,
    nfilters) array from a matrix of either orientation or a list of per-star
    or per-filter vectors."""
    n = len(mini_vec)
    if isinstance(mags, np.ndarray) and mags.ndim == 2:
        if mags.shape[0] == n:
            return np.ascontiguousarray(mags, dtype=np.float64)
        if mags.shape[1] == n:
            return np.ascontiguousarray(mags.T, dtype=np.float64)
        raise ArgumentError("misshapen `mags`: size must be (N, M) or (M, N) with N = length(mini_vec)")
    if isinstance(mags, np.ndarray):
        raise ArgumentError("`mags` must be 2-dimensional when given as a matrix")
    if not isinstance(mags, (list, tuple)):
        raise ArgumentError("there is no `ingest_mags` method for the provided types")
    rows = [np.asarray(v, dtype=np.float64).ravel() for v in mags]
    if len(rows) != n:
        if not all(r.size == n for r in rows):
            raise ArgumentError("each element of `mags` must have length equal to length(mini_vec)")
        return np.ascontiguousarray(np.stack(rows, axis=1))
    if len({r.size for r in rows}) != 1:
        raise ArgumentError("each element of `mags` must have equal length (the number of filters)")
    return np.ascontiguousarray(np.stack(rows, axis=0))


def _nextfloat(x):
    return math.nextafter(x, math.inf)


def sort_ingested(mini_vec, mags):
    """sort_ingested (src/simulate.jl:54-63): masses sorted (magnitudes
    permuted alike), then repeated masses moved up to the next float above
    their predecessor (Interpolations.deduplicate_knots!, move_knots=true)."""
    m = np.asarray(mini_vec, dtype=np.float64)
    g = np.asarray(mags, dtype=np.float64)
    if m.shape[0] != g.shape[0]:
        raise ArgumentError("axes(mini_vec) == axes(mags)")
    idx = np.argsort(m, kind="stable")
    if not np.array_equal(idx, np.arange(m.size)):
        m, g = m[idx], g[idx]
    m = m.copy()
    for i in range(1, m.size):
        if m[i] <= m[i - 1]:
            m[i] = _nextfloat(m[i - 1])
    return m, g


def _bits(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def _from_bits(i):
    return struct.unpack("<d", struct.pack("<q", i))[0]


def _bisect(f, a, b):
    """Bracketing root of f on [a, b] (0 < a < b) by bisection of the float
    representation until f vanishes or the bracket is two adjacent floats."""
    fa, fb = f(a), f(b)
    if fa == 0:
        return a
    if fb == 0:
        return b
    if np.sign(fa) == np.sign(fb):
        raise ArgumentError("the bracket does not contain a sign change")
    while True:
        ia, ib = _bits(a), _bits(b)
        if ib - ia <= 1:
            return a if abs(fa) <= abs(fb) else b
        m = _from_bits((ia + ib) >> 1)
        fm = f(m)
        if fm == 0:
            return m
        if np.sign(fm) == np.sign(fa):
            a, fa = m, fm
        else:
            b, fb = m, fm


def _lininterp(x, xs, ys):
    """Gridded linear interpolation: (1 - d) y_i + d y_{i+1}, d = (x - x_i) / (x_{i+1} - x_i)."""
    i = int(np.clip(np.searchsorted(xs, x, side="right") - 1, 0, xs.size - 2))
    d = (x - xs[i]) / (xs[i + 1] - xs[i])
    return (1.0 - d) * ys[i] + d * ys[i + 1]


def mass_limits(mini_vec, mags, mag_names, mag_lim, mag_lim_name):
    """mass_limits (src/simulate.jl:92-123): (mmin, mmax), mmin raised to the
    initial mass where the y magnitude crosses mag_lim."""
    m = np.asarray(mini_vec, dtype=np.float64)
    g = ingest_mags(m, mags)
    if m.shape[0] != g.shape[0]:
        raise ArgumentError("axes(mini_vec) == axes(mags)")
    mmin, mmax = float(m.min()), float(m.max())
    if not np.isfinite(mag_lim):
        return mmin, mmax
    if mag_lim_name not in list(mag_names):
        raise ArgumentError("Provided `mag_lim_name` is not contained in provided `mag_names` array.")
    idx = list(mag_names).index(mag_lim_name)
    tmp = g[:, idx]
    if mag_lim < tmp[int(np.argmin(m))]:
        if mag_lim < tmp.min():
            raise DomainError("The provided `mag_lim` is brighter than all the stars in the input `mags` array "
                              "assuming the input distance modulus `dist_mod`.")
        brightest = int(np.argmin(tmp))
        root = _bisect(lambda x: _lininterp(x, m, tmp) - mag_lim, mmin, float(m[brightest]))
        return root, mmax
    return mmin, mmax


# ------------------------------------------------------------------ catalogues

class Catalog:
    """A star catalogue resident on the device (sfh_catalog)."""

    def __init__(self, handle, ctx: Context):
        self.handle, self.ctx = handle, ctx
        n, ns, nf, sampled = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
        ctx.check(ctx.lib.sfh_catalog_shape(handle, C.byref(n), C.byref(ns), C.byref(nf), C.byref(sampled)))
        self.n, self.nstars, self.nfilters, self.sampled = n.value, ns.value, nf.value, sampled.value

    @classmethod
    def from_mags(cls, mags, ctx: Context | None = None):
        ctx = ctx or default_context()
        g = np.ascontiguousarray(mags, dtype=np.float64)
        if g.ndim != 2:
            raise ArgumentError("mags must be (nstars, nfilters)")
        h = C.c_void_p()
        ctx.check(ctx.lib.sfh_catalog_from_host(ctx.handle, g.shape[0], g.shape[1], g.ctypes.data, C.byref(h)))
        return cls(h, ctx)

    def arrays(self):
        """(masses (n, nstars) or None, mags (n, nfilters), source indices (n))."""
        m = np.empty((self.n, self.nstars)) if self.nstars else None
        g = np.empty((self.n, self.nfilters))
        ix = np.empty(self.n, dtype=np.int64)
        self.ctx.check(self.ctx.lib.sfh_catalog_copy(self.handle, m.ctypes.data if m is not None else None,
                                                     g.ctypes.data, ix.ctypes.data))
        return m, g, ix

    def mag_range(self):
        """Per-filter (min, max) of the finite magnitudes, reduced on the device."""
        lo, hi = np.empty(self.nfilters), np.empty(self.nfilters)
        self.ctx.check(self.ctx.lib.sfh_catalog_range(self.handle, lo.ctypes.data, hi.ctypes.data))
        return lo, hi

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.sfh_catalog_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _generate(mini_vec, mags, mag_names, mode, limit, limit_filter, imf, dist_mod, seed, stream, mag_lim,
              mag_lim_name, binary_model, ctx):
    ctx = ctx or default_context()
    if not isinstance(imf, PowerLawIMF):
        raise ArgumentError("imf must be a PowerLawIMF (Salpeter1955, Kroupa2001 or PowerLawIMF)")
    g = ingest_mags(mini_vec, mags) + float(dist_mod)                 # :309
    m, g = sort_ingested(mini_vec, g)                                  # :310
    if g.shape[1] > _lib.MOCK_MAX_FILTERS:
        raise ArgumentError(f"at most {_lib.MOCK_MAX_FILTERS} filters")
    names = list(mag_names)
    mmin_cut = -math.inf
    mlf = -1
    if mode == 0:
        mmin_cut, _ = mass_limits(m, g, names, mag_lim, mag_lim_name)  # :315
    elif np.isfinite(mag_lim):
        if mag_lim_name not in names:
            raise ArgumentError("Provided `mag_lim_name` is not contained in provided `mag_names` array.")
        mlf = names.index(mag_lim_name)
    bm, _ = _binary_c(binary_model)
    m = np.ascontiguousarray(m)
    g = np.ascontiguousarray(g)
    h = C.c_void_p()
    try:
        ctx.check(ctx.lib.sfh_generate_stars(ctx.handle, m.size, m.ctypes.data, g.shape[1], g.ctypes.data, mode,
                                             float(limit), int(limit_filter), float(mmin_cut), float(mag_lim), mlf,
                                             C.byref(imf.as_c()), C.byref(bm), int(seed) & (2 ** 64 - 1),
                                             int(stream) & (2 ** 64 - 1), C.byref(h)))
    except _lib.InexactError as e:          # the library's domain errors: an unreachable limit
        raise DomainError(str(e)) from None
    return Catalog(h, ctx)


def generate_stars_mass(mini_vec, mags, mag_names, limit, imf, *, dist_mod=0.0, seed=0, stream=0,
                        mag_lim=math.inf, mag_lim_name="V", binary_model=None, catalog=False, ctx=None):
    """generate_stars_mass (src/simulate.jl:306-339): systems drawn until their
    birth masses reach ``limit``.  Returns (masses (n, nstars), mags (n,
    nfilters)), or the device Catalog with catalog=True."""
    bm = RandomBinaryPairs(0.3) if binary_model is None else binary_model
    c = _generate(mini_vec, mags, mag_names, 0, limit, 0, imf, dist_mod, seed, stream, mag_lim, mag_lim_name, bm,
                  ctx)
    if catalog:
        return c
    masses, g, _ = c.arrays()
    c.close()
    return masses, g


def generate_stars_mag(mini_vec, mags, mag_names, absmag, absmag_name, imf, *, dist_mod=0.0, seed=0, stream=0,
                       mag_lim=math.inf, mag_lim_name="V", binary_model=None, catalog=False, ctx=None):
    """generate_stars_mag (src/simulate.jl:350-404): systems drawn until their
    flux in ``absmag_name`` reaches that of absolute magnitude ``absmag``."""
    names = list(mag_names)
    if absmag_name not in names:
        raise ArgumentError("Provided `absmag_name` is not contained in provided `mag_names` array.")
    bm = RandomBinaryPairs(0.3) if binary_model is None else binary_model
    limit = float(mag2flux(absmag + dist_mod))                         # :359
    c = _generate(mini_vec, mags, mag_names, 1, limit, names.index(absmag_name), imf, dist_mod, seed, stream,
                  mag_lim, mag_lim_name, bm, ctx)
    if catalog:
        return c
    masses, g, _ = c.arrays()
    c.close()
    return masses, g


def generate_stars_mass_composite(mini_vec, mags, mag_names, limit, massfrac, imf, *, binary_model=None, seed=0,
                                  **kws):
    """generate_stars_mass_composite (src/simulate.jl:428-447): ``limit``
    split by ``massfrac`` (normalised) over the isochrones, population i
    drawn with stream i.  Returns (list of masses, list of mags)."""
    if not (len(mini_vec) == len(mags) == len(massfrac)):
        raise ArgumentError("The arguments `mini_vec`, `mags`, and `massfrac` to `generate_stars_mass_composite` "
                            "must all have equal length and identical indexing.")
    f = np.asarray(massfrac, dtype=np.float64)
    f = f / f.sum()
    out_m, out_g = [], []
    for i in range(len(mini_vec)):
        a, b = generate_stars_mass(mini_vec[i], mags[i], mag_names, limit * f[i], imf, binary_model=binary_model,
                                   seed=seed, stream=i, **kws)
        out_m.append(a)
        out_g.append(b)
    return out_m, out_g


def generate_stars_mag_composite(mini_vec, mags, mag_names, absmag, absmag_name, fracs, imf, *, frac_type="lum",
                                 binary_model=None, seed=0, **kws):
    """generate_stars_mag_composite (src/simulate.jl:473-499): the luminosity
    of ``absmag`` split by ``fracs`` (normalised) over the isochrones."""
    if not (len(mini_vec) == len(mags) == len(fracs)):
        raise ArgumentError("The arguments `mini_vec`, `mags`, and `fracs` to `generate_stars_mag_composite` must "
                            "all have equal length and identical indexing.")
    if frac_type == "mass":
        raise ArgumentError("`frac_type == :mass` not yet implemented.")
    if frac_type != "lum":
        raise ArgumentError("Supported `frac_type` arguments for generate_stars_mag_composite are `:lum` or `:mass`.")
    f = np.asarray(fracs, dtype=np.float64)
    f = f / f.sum()
    absmags = flux2mag(f * mag2flux(absmag))                           # :483-484
    out_m, out_g = [], []
    for i in range(len(mini_vec)):
        a, b = generate_stars_mag(mini_vec[i], mags[i], mag_names, float(absmags[i]), absmag_name, imf,
                                  binary_model=binary_model, seed=seed, stream=i, **kws)
        out_m.append(a)
        out_g.append(b)
    return out_m, out_g


# ------------------------------------------------------------------ observing

def _tables(funcs, lo, hi):
    """Tabulate callables over [lo, hi] (one per filter): sfh_table array and
    the arrays backing it."""
    tabs = (_lib.sfh_table * len(funcs))()
    keep = []
    for f, fn in enumerate(funcs):
        a, b = float(lo[f]), float(hi[f])
        if not b > a:
            a, b = a - 0.5, b + 0.5
        grid = np.linspace(a, b, _TABLE_POINTS)
        vals = np.ascontiguousarray(_apply(fn, grid))
        keep.append(vals)
        tabs[f] = _lib.sfh_table(float(grid[0]), float(grid[1] - grid[0]), vals.size, vals.ctypes.data)
    return tabs, keep


def model_cmd(mags, errfuncs, completefuncs, biasfuncs=None, *, seed=0, stream=0, ret_idxs=False, ctx=None):
    """model_cmd (src/simulate.jl:621-674): keeps star j with probability the
    product of completefuncs over filters, then adds Gaussian error
    errfuncs[f](mag) and bias biasfuncs[f](mag).  ``mags``: (nstars, nfilters)
    array or a device Catalog.  Returns new mags (and the kept input indices
    with ret_idxs=True), or a Catalog when given one."""
    ctx = ctx or (mags.ctx if isinstance(mags, Catalog) else default_context())
    if isinstance(mags, Catalog):
        cat, own = mags, False
    else:
        host = np.ascontiguousarray(mags, dtype=np.float64)
        if host.ndim != 2:
            raise ArgumentError("mags must be (nstars, nfilters)")
        cat, own = Catalog.from_mags(host, ctx), True
    nf = cat.nfilters
    if not (nf == len(errfuncs) == len(completefuncs)) or (biasfuncs is not None and len(biasfuncs) != nf):
        if own:
            cat.close()
        raise ArgumentError("Arguments to `StarFormationHistories.model_cmd` must satisfy "
                            "`length(first(mags)) == length(errfuncs) == length(completefuncs)`.")
    lo, hi = cat.mag_range()
    empty = ~(np.isfinite(lo) & np.isfinite(hi))
    lo, hi = np.where(empty, 0.0, lo), np.where(empty, 1.0, hi)
    et, k1 = _tables(errfuncs, lo, hi)
    ct, k2 = _tables(completefuncs, lo, hi)
    bt, k3 = _tables(biasfuncs, lo, hi) if biasfuncs is not None else (None, [])
    h = C.c_void_p()
    ctx.check(ctx.lib.sfh_model_cmd(ctx.handle, cat.handle, et, ct, bt, int(seed) & (2 ** 64 - 1),
                                    int(stream) & (2 ** 64 - 1), C.byref(h)))
    out = Catalog(h, ctx)
    del k1, k2, k3
    if own:
        cat.close()
    if isinstance(mags, Catalog):
        return out
    _, g, ix = out.arrays()
    out.close()
    return (g, ix) if ret_idxs else g


def bin_cmd(colors, mags, *, weights=None, edges=None, xlim=None, ylim=None, nbins=None, xwidth=None, ywidth=None,
            ctx=None) -> Histogram:
    """bin_cmd (src/StarFormationHistories.jl:544-553): the left-closed
    weighted Hess diagram of (colors, mags)."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(colors, dtype=np.float64).ravel()
    y = np.ascontiguousarray(mags, dtype=np.float64).ravel()
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
    if not (x.size == y.size and (w is None or w.size == x.size)):
        raise ArgumentError("length(colors) == length(mags) == length(weights)")
    if edges is None and xlim is None:
        xlim = (float(x.min()), float(x.max()))
    if edges is None and ylim is None:
        ylim = (float(y.min()), float(y.max()))
    ex, ey = calculate_edges(edges, xlim, ylim, nbins, xwidth, ywidth)
    nx, ny = ex.length - 1, ey.length - 1
    out = np.empty(nx * ny)
    ctx.check(ctx.lib.sfh_bin_cmd(ctx.handle, x.size, x.ctypes.data, y.ctypes.data,
                                  w.ctypes.data if w is not None else None, ex.as_c(), ey.as_c(), out.ctypes.data))
    return Histogram((ex, ey), out.reshape((nx, ny), order="F"))


def bin_cmd_catalog(cat: Catalog, color_filters, y_filter, edges) -> Histogram:
    """bin_cmd of a device catalogue: colour = filter a - filter b, y = y_filter."""
    ex, ey = edges
    nx, ny = ex.length - 1, ey.length - 1
    out = np.empty(nx * ny)
    cat.ctx.check(cat.ctx.lib.sfh_bin_cmd_catalog(cat.ctx.handle, cat.handle, int(color_filters[0]),
                                                  int(color_filters[1]), int(y_filter), ex.as_c(), ey.as_c(),
                                                  out.ctypes.data))
    return Histogram((ex, ey), out.reshape((nx, ny), order="F"))


# ------------------------------------------------------------------ population bookkeeping

def surviving_fraction(massvec, coeffs, imf):
    """surviving_fraction (src/simulate.jl:684-693)."""
    if len(massvec) != len(coeffs):
        raise ArgumentError("Length of `massvec` and `coeffs` must be equal.")
    tot = float(np.sum(coeffs))
    s = 0.0
    for m, c in zip(massvec, coeffs):
        lo, hi = float(np.min(m)), float(np.max(m))
        s += c * float(imf.cdf(hi)[0] - imf.cdf(lo)[0])
    return s / tot


def recycling_fraction(massvec, coeffs, imf):
    """recycling_fraction (src/simulate.jl:699)."""
    return 1.0 - surviving_fraction(massvec, coeffs, imf)


def surviving_mass_fraction(massvec, coeffs, imf):
    """surviving_mass_fraction (src/simulate.jl:704-708)."""
    if len(massvec) != len(coeffs):
        raise ArgumentError("Length of `massvec` and `coeffs` must be equal.")
    s = sum(trapz(m, np.asarray(m) * imf.pdf(m)) * c for m, c in zip(massvec, coeffs))
    return s / float(np.sum(coeffs)) / imf.mean()


def recycling_mass_fraction(massvec, coeffs, imf):
    """recycling_mass_fraction (src/simulate.jl:713)."""
    return 1.0 - surviving_mass_fraction(massvec, coeffs, imf)


__all__ = ["PowerLawIMF", "Salpeter1955", "Kroupa2001", "BinaryMassRatio", "Catalog", "ingest_mags",
           "sort_ingested", "mass_limits", "generate_stars_mass", "generate_stars_mag",
           "generate_stars_mass_composite", "generate_stars_mag_composite", "model_cmd", "bin_cmd",
           "bin_cmd_catalog", "surviving_fraction", "recycling_fraction", "surviving_mass_fraction",
           "recycling_mass_fraction", "Range"]
