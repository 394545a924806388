"""Seeded synthetic inputs for the template and likelihood paths.

No isochrone tables or datasets are available offline, so every workload is
generated here from a seed: smooth colour-magnitude curves parameterised by
(logAge, [M/H]), a Kroupa (2001) broken power-law IMF, the photometric error
sigma(m) = 0.01 + exp(0.6 (m - 28)) and a logistic completeness (50% at
m = 28, slope 0.7 mag as in the reference's own template test,
test/templates/template_test.jl:72-74).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .templates import (ExpPhotErr, Martin2016Completeness, Martin2016_complete, Range, edges_range, exp_photerr,
                        partial_cmd_smooth_batch)


class Kroupa2001:
    """Kroupa (2001) IMF, dN/dm ∝ m^-1.3 on [0.08, 0.5) and ∝ 0.5 m^-2.3 above,
    normalised on [mmin, mmax] (mmin >= 0.08).  Restated from the published
    form; the reference's tests use InitialMassFunctions.Kroupa2001, which is
    not vendored."""

    def __init__(self, mmin=0.08, mmax=100.0):
        if mmin < 0.08:
            raise ValueError("mmin >= 0.08 only")
        self.mmin, self.mmax = float(mmin), float(mmax)
        self._norm = 1.0 / self._integral(lambda a, lo, hi: self._pow_int(a, lo, hi))

    @staticmethod
    def _pow_int(alpha, lo, hi):      # ∫ m^-alpha dm
        if lo >= hi:
            return 0.0
        p = 1.0 - alpha
        return (hi ** p - lo ** p) / p

    def _integral(self, f, shift=0.0):
        lo, hi, br = self.mmin, self.mmax, 0.5
        s = 0.0
        s += f(1.3 - shift, lo, min(hi, br))
        s += 0.5 * f(2.3 - shift, max(lo, br), hi)
        return s

    def __call__(self, m):
        m = np.asarray(m, dtype=np.float64)
        out = np.where(m < 0.5, m ** -1.3, 0.5 * m ** -2.3) * self._norm
        return np.where((m >= self.mmin) & (m <= self.mmax), out, 0.0)

    def mean(self):
        return self._norm * self._integral(lambda a, lo, hi: self._pow_int(a, lo, hi), shift=1.0)

    def as_fn(self, lo=None, hi=None):
        """The device form: normalisation x coefs[k] m^-slopes[k] on the
        segments of [mmin, mmax] split at 0.5 Msun."""
        if self.mmin < 0.5 < self.mmax:
            br, sl, co = [self.mmin, 0.5, self.mmax], [1.3, 2.3], [1.0, 0.5]
        elif self.mmax <= 0.5:
            br, sl, co = [self.mmin, self.mmax], [1.3], [1.0]
        else:
            br, sl, co = [self.mmin, self.mmax], [2.3], [0.5]
        self._c = [np.ascontiguousarray(v, dtype=np.float64) for v in (br, sl, co)]
        fn = _lib.sfh_fn(_lib.SFH_FN_POWERLAW)
        fn.p[0] = self._norm
        fn.powerlaw = _lib.sfh_imf(len(sl), *(v.ctypes.data for v in self._c))
        return fn


# the same error and completeness models as objects the device evaluates exactly
PHOTERR = ExpPhotErr(math.e, 0.6, 28.0, 0.01)
COMPLETENESS = Martin2016Completeness(1.0, 28.0, 0.7)


def photerr(m):
    """sigma(m) = 0.01 + exp(0.6 (m - 28)) = exp_photerr(m, e, 0.6, 28, 0.01)."""
    return exp_photerr(m, math.e, 0.6, 28.0, 0.01)


def completeness(m, rho=0.7):
    """Logistic completeness, 50% at m = 28 (Martin2016_complete(m, 1, 28, rho))."""
    return Martin2016_complete(m, 1.0, 28.0, rho)


def turnoff_mass(log_age):
    return (10.0 ** (10.0 - log_age)) ** 0.4


def isochrone(log_age, mh, npoints, seed, mmax_imf=120.0):
    """A smooth synthetic isochrone: initial masses (increasing) and absolute
    magnitudes in two filters [B, V] such that V + 24 spans ~18-31 for the
    main sequence and colour B-V lies in about -0.4..2.3.  Main sequence up to
    the turnoff mass, then a giant branch reddening and brightening over the
    last 8% in mass.  ``seed`` adds small smooth wiggles (<~0.02 mag)."""
    rng = np.random.default_rng(seed)
    mto = turnoff_mass(log_age)
    mend = min(1.08 * mto, mmax_imf)
    m = np.exp(np.linspace(math.log(0.08), math.log(mend), npoints))
    m[-1] = mend
    lm = np.log10(m / 0.08)
    V = 6.5 - 5.93 * lm                        # 0.08 Msun -> 6.5, 1 Msun -> ~0
    col = 2.2 - 1.25 * lm + 0.18 * mh
    col = -0.4 + np.logaddexp(0.0, 4.0 * (col + 0.4)) / 4.0     # smooth floor at -0.4
    gb = m > mto
    if np.any(gb):
        x = (m[gb] - mto) / (mend - mto)
        Vto = 6.5 - 5.93 * math.log10(mto / 0.08)
        cto = np.interp(mto, m, col)
        V[gb] = Vto - 3.0 * x ** 1.5
        col[gb] = cto + (1.5 + 0.25 * mh - cto) * np.sqrt(x)
    s = np.linspace(0.0, 1.0, npoints)
    for k in range(1, 4):
        a, ph = rng.uniform(0.003, 0.012, 2), rng.uniform(0, 2 * np.pi, 2)
        V += a[0] * np.sin(2 * np.pi * k * s + ph[0])
        col += a[1] * np.sin(2 * np.pi * k * s + ph[1])
    # the last segment is kept shorter than any bin so mini_spacing reproduces
    # the final mass exactly (no 1-ulp overshoot past the interpolation knots)
    V[-1] = V[-2] + 1e-4
    col[-1] = col[-2]
    B = V + col
    return m, [B, V]


@dataclass
class TemplateSpec:
    log_age: float
    mh: float
    seed: int


def template_grid(log_ages, mhs, base_seed):
    specs = []
    for i, la in enumerate(log_ages):
        for j, mh in enumerate(mhs):
            specs.append(TemplateSpec(float(la), float(mh), int(base_seed * 1000003 + i * 1009 + j)))
    return specs


def edges_75x100():
    """75 x 100 bins (76 / 101 edges) over colour -0.5..2.5, mag 18..30."""
    return (edges_range(-0.5, 2.5, 76), edges_range(18.0, 30.0, 101))


def edges_200x300():
    """200 x 300 bins (201 / 301 edges), 60,000 bins."""
    return (edges_range(-0.5, 2.5, 201), edges_range(18.0, 30.0, 301))


def specs_10x5(seed=20261202):
    return template_grid(np.linspace(8.0, 10.1, 10), np.linspace(-2.0, 0.0, 5), seed)


def specs_70x10(seed=20261202, n_age=70, n_mh=10):
    return template_grid(6.6 + 0.05 * np.arange(n_age), np.linspace(-2.5, 0.0, n_mh), seed + 1)


def template_points(spec: TemplateSpec, npoints, edges, mmax_imf, dmod=24.0, normalize_value=1.0,
                    correlation=None, err_scale=1.0):
    """Host preparation (partial_cmd_smooth_points) of one synthetic template.
    y = V, colour = B - V (cov_mult = +1).  With ``correlation`` set, also
    returns per-point covariance multipliers c = -rho/sqrt(1-rho^2) * sx/sy,
    which give the covariant kernel exactly that colour-magnitude correlation
    (corr = -c sy / sqrt(sx^2 + c^2 sy^2)).  ``err_scale`` multiplies the
    photometric error model."""
    from .templates import partial_cmd_smooth_points
    imf = Kroupa2001(0.08, mmax_imf)
    m, mags = isochrone(spec.log_age, spec.mh, npoints, spec.seed, mmax_imf=mmax_imf)
    err = photerr if err_scale == 1.0 else (lambda v, s=float(err_scale): s * photerr(v))
    sp = partial_cmd_smooth_points(m, mags, [err, err], 1, (0, 1), imf,
                                   [completeness, completeness], None, dmod=dmod,
                                   normalize_value=normalize_value, mean_mass=imf.mean(), edges=edges)
    cm = None
    if correlation is not None:
        r = float(correlation)
        cm = -r / math.sqrt(1.0 - r * r) * sp.color_err / sp.mag_err
    return sp, cm


def templates_device(specs, npoints, edges, mmax_imf, dmod=24.0, normalize_value=1.0, correlation=0.0,
                     output="device", return_info=False, ctx=None):
    """The isochrones of ``specs`` built by partial_cmd_smooth_batch (the
    preparation on the device): y = V, colour = B - V, the same error,
    completeness and IMF as template_points.  The synthetic isochrones are
    generated on the host first (inputs, not part of the template path)."""
    imf = Kroupa2001(0.08, mmax_imf)
    isos = [isochrone(sp.log_age, sp.mh, npoints, sp.seed, mmax_imf=mmax_imf) for sp in specs]
    return partial_cmd_smooth_batch(isos, [PHOTERR, PHOTERR], 1, (0, 1), imf, [COMPLETENESS, COMPLETENESS],
                                    dmod=dmod, normalize_value=normalize_value, mean_mass=imf.mean(),
                                    correlation=correlation, edges=edges, output=output, return_info=return_info,
                                    ctx=ctx)


def poisson_data(composite, seed):
    rng = np.random.default_rng(seed)
    return rng.poisson(np.maximum(composite, 0.0)).astype(np.float64)


def lognormal_coeffs(T, total, sigma_dex, seed):
    rng = np.random.default_rng(seed)
    x = 10.0 ** (sigma_dex * rng.standard_normal(T))
    return x


def synthetic_templates(nbins, ntemplates, seed, dtype=np.float32):
    """Seeded FP32 templates of a given shape (bands of smooth texture, unit
    sum each), a stand-in for deposited templates when only the likelihood
    path is exercised."""
    rng = np.random.default_rng(seed)
    i = np.arange(nbins, dtype=np.float64)
    out = np.empty((nbins, ntemplates), dtype=dtype, order="F")
    for j in range(ntemplates):
        c = rng.uniform(0, nbins)
        w = rng.uniform(0.02, 0.3) * nbins
        tex = np.exp(0.3 * np.sin(i * rng.uniform(0.001, 0.05) + rng.uniform(0, 6.28)))
        col = np.exp(-0.5 * ((i - c) / w) ** 2) * tex + 1e-6
        out[:, j] = col / col.sum()
    return out


def walkers(x_true, B, sigma_dex, seed):
    """B coefficient vectors log-normal around x_true (sigma in dex)."""
    rng = np.random.default_rng(seed)
    T = x_true.size
    return np.asfortranarray(x_true[:, None] * 10.0 ** (sigma_dex * rng.standard_normal((T, B))))
