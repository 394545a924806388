"""CPU restatement (NumPy, FP64) of the mock-observation path, for tests and
the bench's host baseline only: generate_stars_mass / generate_stars_mag with
sample_system (src/simulate.jl:246-404), model_cmd (:621-674) and bin_cmd
(src/StarFormationHistories.jl:544-553), drawing the same counter-based random
numbers as the GPU kernels (Philox4x64-10, one block per system or star), so
that the GPU catalogues can be compared entry by entry.

Pins (tests/test_oracle_mock.py): philox4x64 against numpy.random.Philox
(bit-exact, an independent implementation); the IMF sampler against its
analytic CDF (Kolmogorov distance) and mean; the fixed-point stopping rule
against the reference's sequential FP64 loop; the reference's own
statistical checks of generate_stars_mass / generate_stars_mag
(test/cmd_simulation_test.jl:133-214) on the reference's test isochrone;
model_cmd against its closed-form limits (unit completeness and zero error
give the input back; kept fraction and error spread binomial / normal);
bin_cmd against numpy.histogram2d.
"""
from __future__ import annotations

import math

import numpy as np

M32 = np.uint64(0xFFFFFFFF)
_M0, _M1 = np.uint64(0xD2E7470EE14C6C93), np.uint64(0xCA5A826395121157)
_W0, _W1 = np.uint64(0x9E3779B97F4A7C15), np.uint64(0xBB67AE8584CAA73B)
TARGET = 1 << 62


def _mulhilo(a, m):
    """(hi, lo) of the 128-bit product of uint64 arrays a and constant m."""
    a_lo, a_hi = a & M32, a >> np.uint64(32)
    m_lo, m_hi = m & M32, m >> np.uint64(32)
    p0, p1, p2, p3 = a_lo * m_lo, a_lo * m_hi, a_hi * m_lo, a_hi * m_hi
    mid = (p0 >> np.uint64(32)) + (p1 & M32) + (p2 & M32)
    lo = a * m
    hi = p3 + (p1 >> np.uint64(32)) + (p2 >> np.uint64(32)) + (mid >> np.uint64(32))
    return hi, lo


def philox4x64(c0, c1, k0, k1):
    """Philox4x64-10 (Salmon et al. 2011) of counters (c0, c1, 0, 0) and key
    (k0, k1); c0 / c1 arrays or scalars.  Returns four uint64 arrays."""
    with np.errstate(over="ignore"):
        x0 = np.asarray(c0, dtype=np.uint64).copy()
        x1 = np.broadcast_to(np.asarray(c1, dtype=np.uint64), x0.shape).copy()
        x2 = np.zeros_like(x0)
        x3 = np.zeros_like(x0)
        k0, k1 = np.uint64(k0), np.uint64(k1)
        for _ in range(10):
            hi0, lo0 = _mulhilo(x0, _M0)
            hi1, lo1 = _mulhilo(x2, _M1)
            x0, x1, x2, x3 = hi1 ^ x1 ^ k0, lo1, hi0 ^ x3 ^ k1, lo0
            k0 = k0 + _W0
            k1 = k1 + _W1
    return x0, x1, x2, x3


def u01(r):
    return (np.asarray(r, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


# ----------------------------------------------------------------- the IMF

class PowerLaw:
    """pdf(m) = coef_k m^-slope_k on [breaks[k], breaks[k+1]) (unnormalised)."""

    def __init__(self, breaks, slopes, coefs):
        self.lo = np.asarray(breaks, dtype=np.float64)
        self.a = 1.0 - np.asarray(slopes, dtype=np.float64)
        self.coef = np.asarray(coefs, dtype=np.float64)
        self.nseg = self.a.size
        cum = [0.0]
        for k in range(self.nseg):
            lo, hi, a, c = self.lo[k], self.lo[k + 1], self.a[k], self.coef[k]
            cum.append(cum[-1] + (c * math.log(hi / lo) if a == 0 else c * (hi ** a - lo ** a) / a))
        self.cum = np.array(cum)

    def sample(self, u):
        x = np.asarray(u, dtype=np.float64) * self.cum[-1]
        k = np.zeros(x.shape, dtype=np.int64)
        for j in range(1, self.nseg):
            k = np.where(x >= self.cum[j], j, k)
        y = (x - self.cum[k]) / self.coef[k]
        a, lo = self.a[k], self.lo[k]
        with np.errstate(divide="ignore", invalid="ignore"):
            m = np.where(a == 0, lo * np.exp(y), np.power(np.power(lo, a) + y * a, 1.0 / np.where(a == 0, 1.0, a)))
        return np.minimum(np.maximum(m, lo), self.lo[k + 1])

    def cdf(self, m):
        m = np.asarray(m, dtype=np.float64)
        out = np.zeros(m.shape)
        for k in range(self.nseg):
            lo, hi, a, c = self.lo[k], self.lo[k + 1], self.a[k], self.coef[k]
            mm = np.clip(m, lo, hi)
            part = c * np.log(mm / lo) if a == 0 else c * (mm ** a - lo ** a) / a
            out += np.where(m > lo, part, 0.0)
        return out / self.cum[-1]

    def mean(self):
        s = 0.0
        for k in range(self.nseg):
            lo, hi, a2, c = self.lo[k], self.lo[k + 1], self.a[k] + 1.0, self.coef[k]
            s += c * math.log(hi / lo) if a2 == 0 else c * (hi ** a2 - lo ** a2) / a2
        return s / self.cum[-1]


# ----------------------------------------------------------------- sampling

def _interp(mini, mags, m):
    """Linear interpolation of every filter at masses m (in range)."""
    k = np.clip(np.searchsorted(mini, m, side="right") - 1, 0, mini.size - 2)
    t = (m - mini[k]) / (mini[k + 1] - mini[k])
    y0, y1 = mags[k], mags[k + 1]
    return y0 + t[:, None] * (y1 - y0)


def systems(i, imf: PowerLaw, bkind, bfrac, qmin, qmax, key0, key1):
    """sample_system (src/simulate.jl:246-277) for systems i: masses (n, nst)."""
    r0, r1, r2, _ = philox4x64(i, 0, key0, key1)
    ub, u1, u2 = u01(r0), u01(r1), u01(r2)
    if bkind == 0:
        return imf.sample(u1)[:, None]
    if bkind == 1:
        m1 = imf.sample(u1)
        m2 = np.where(ub <= bfrac, imf.sample(u2), 0.0)
        return np.stack([np.maximum(m1, m2), np.where(m2 > m1, m1, m2)], axis=1)
    M = imf.sample(u1)
    q = qmin + (qmax - qmin) * u2
    b = ub <= bfrac
    p = np.where(b, M / (q + 1.0), M)
    return np.stack([p, np.where(b, M - p, 0.0)], axis=1)


def evaluate(ms, mini, mags, mode):
    """Summed flux per filter of each system's valid stars (mass mode: open
    interval, luminosity mode: closed -- src/simulate.jl:327, :387)."""
    nf = mags.shape[1]
    flux = np.zeros((ms.shape[0], nf))
    for s in range(ms.shape[1]):
        m = ms[:, s]
        valid = (m > mini[0]) & (m < mini[-1]) if mode == 0 else ~((m < mini[0]) | (m > mini[-1]))
        if np.any(valid):
            g = _interp(mini, mags, m[valid])
            flux[valid] += 10.0 ** (4.0 * (0.0 - g) / 10.0)
    return flux


def generate(mini, mags, mode, limit, imf: PowerLaw, bkind=0, bfrac=0.0, qmin=0.1, qmax=1.0, *, limit_filter=0,
             mmin_cut=-np.inf, mag_lim=np.inf, mag_lim_filter=0, seed=0, stream=0, chunk=1 << 18):
    """The GPU's sampling rule: systems 0..K with K the first index at which
    the running fixed-point total (contribution * 2^62 / limit, rounded,
    clamped at 2^62) reaches 2^62.  Returns (masses (n, nst), mags (n, nf),
    system indices, systems drawn)."""
    mini = np.asarray(mini, dtype=np.float64)
    mags = np.asarray(mags, dtype=np.float64)
    scale = 2.0 ** 62 / limit
    total = 0
    out_m, out_g, out_i = [], [], []
    s0 = 0
    while True:
        i = np.arange(s0, s0 + chunk, dtype=np.uint64)
        ms = systems(i, imf, bkind, bfrac, qmin, qmax, seed, stream)
        flux = evaluate(ms, mini, mags, mode)
        fsum = flux.sum(axis=1)
        with np.errstate(divide="ignore"):
            g = 0.0 - 5.0 * np.log10(flux) / 2.0
        if mode == 0:
            c = ms.sum(axis=1)
            keep = ~(ms[:, 0] < mmin_cut) & (fsum > 0)
        else:
            c = flux[:, limit_filter]
            keep = fsum > 0
            if np.isfinite(mag_lim):
                keep &= g[:, mag_lim_filter] < mag_lim
        qd = c * scale
        q = np.rint(np.minimum(qd, float(TARGET))).astype(np.uint64)      # clamped at 2^62
        cs = np.uint64(total) + np.cumsum(q, dtype=np.uint64)
        hit = np.flatnonzero(cs >= np.uint64(TARGET))
        kend = int(hit[0]) + 1 if hit.size else chunk
        keep = keep[:kend]
        out_m.append(ms[:kend][keep])
        out_g.append(g[:kend][keep])
        out_i.append(np.arange(s0, s0 + kend)[keep])
        s0 += kend
        if hit.size:
            break
        total = int(cs[-1])
    return np.concatenate(out_m), np.concatenate(out_g), np.concatenate(out_i), s0


def sequential_stop(contrib, limit):
    """The reference's loop (src/simulate.jl:322-324): FP64 running total,
    stop after the system that brings it to the limit."""
    total = 0.0
    for k, c in enumerate(contrib):
        total += float(c)
        if not total < limit:
            return k + 1
    return None


# ----------------------------------------------------------------- observing

def table_eval(first, step, values, x):
    """Linear interpolation on a uniform grid, clamped to the ends."""
    v = np.asarray(values, dtype=np.float64)
    u = np.clip((np.asarray(x, dtype=np.float64) - first) / step, 0.0, v.size - 1)
    k = np.minimum(u.astype(np.int64), v.size - 2)
    f = u - k
    return v[k] + f * (v[k + 1] - v[k])


def normals(j, nf, key0, key1):
    """Standard normals (n, nf) of stars j: Box-Muller pair f // 2 (pair 0 from
    words 1, 2 of block 0, pair p > 0 from words 0, 1 of block p); cosine for
    even f, sine for odd f."""
    r = philox4x64(j, 0, key0, key1)
    out = np.empty((np.size(j), nf))
    for f in range(nf):
        p = f // 2
        if p == 0:
            wa, wb = r[1], r[2]
        else:
            rp = philox4x64(j, p, key0, key1)
            wa, wb = rp[0], rp[1]
        rad = np.sqrt(-2.0 * np.log(1.0 - u01(wa)))
        ang = 2.0 * np.pi * u01(wb)
        out[:, f] = rad * (np.sin(ang) if f & 1 else np.cos(ang))
    return out


def model_cmd(mags, err_tables, compl_tables, bias_tables=None, seed=0, stream=0):
    """model_cmd (src/simulate.jl:621-674) with tabulated callables (first,
    step, values) per filter.  Returns (new mags, kept input indices)."""
    mags = np.asarray(mags, dtype=np.float64)
    n, nf = mags.shape
    j = np.arange(n, dtype=np.uint64)
    r0 = philox4x64(j, 0, seed, stream)[0]
    c = np.ones(n)
    for f in range(nf):
        c = c * table_eval(*compl_tables[f], mags[:, f])
    keep = u01(r0) <= c
    z = normals(j, nf, seed, stream)
    out = np.empty_like(mags)
    for f in range(nf):
        e = table_eval(*err_tables[f], mags[:, f])
        b = table_eval(*bias_tables[f], mags[:, f]) if bias_tables is not None else 0.0
        out[:, f] = mags[:, f] + z[:, f] * e + b
    idx = np.flatnonzero(keep)
    return out[idx], idx


def bin_cmd(colors, mags, x0, dx, nx, y0, dy, ny, weights=None):
    """Left-closed 2-D histogram (nx, ny) with bins floor((v - first) / step)."""
    fx = np.floor((np.asarray(colors, dtype=np.float64) - x0) / dx)
    fy = np.floor((np.asarray(mags, dtype=np.float64) - y0) / dy)
    ok = (fx >= 0) & (fx < nx) & (fy >= 0) & (fy < ny)
    w = np.ones(fx.size) if weights is None else np.asarray(weights, dtype=np.float64)
    h = np.zeros(nx * ny)
    np.add.at(h, fx[ok].astype(np.int64) + nx * fy[ok].astype(np.int64), w[ok])
    return h.reshape(ny, nx).T
