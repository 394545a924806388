"""Host-side callers of the GPU likelihood: the reference's optimisers and
sampler models, unchanged in meaning, with every objective evaluation on the
GPU (fg_ / fg_batched).

Mirrors src/fitting/solvers.jl (fit_templates, fit_templates_fast,
fit_templates_lbfgsb), src/fitting/utilities.jl (renormalize_x0,
construct_x0), src/fitting/hmc_sample.jl (HMCModel) and
src/fitting/mcmc_sample.jl (MCMCModel, mcmc_sample's affine-invariant
ensemble).  The reference drives Optim.jl's BFGS with a Hager-Zhang line
search and LBFGSB.jl; here SciPy's BFGS (Wolfe line search) and L-BFGS-B
(the same Nocedal-Zhu Fortran algorithm) play those roles, so iterates differ
while the optimum they converge to is the same.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import ArgumentError
from .core import HIPData, HIPTemplates, composite_, driver_clamps, fg_, fg_batched, loglikelihood, stack_models


def _models_data(models, data):
    tpl = models if isinstance(models, HIPTemplates) else HIPTemplates(models)
    hd = data if isinstance(data, HIPData) else HIPData(np.asarray(data, dtype=np.float64).ravel(order="F"),
                                                         ctx=tpl.ctx)
    if hd.nbins != tpl.nbins:
        raise ArgumentError("axes(models,1) == axes(data,1)")
    return tpl, hd


def _data_sum(data) -> float:
    if isinstance(data, HIPData):
        raise ArgumentError("renormalize_x0 needs host data")
    return float(np.sum(np.asarray(data, dtype=np.float64)))


def renormalize_x0(data, models, x0, full_coeffs=None):
    """renormalize_x0 (src/fitting/utilities.jl:104-115): scale x0 by
    sum(data) / sum(composite(full_coeffs)).  ``models`` is a HIPTemplates, a
    flattened nbins x ntemplates matrix, or a list of 2-D templates with
    ``data`` a matrix (the vector-of-matrices method, :117-122)."""
    if not isinstance(models, HIPTemplates):
        if isinstance(models, (list, tuple)):
            models = stack_models(models)
            data = np.asarray(data).ravel(order="F")
        models = HIPTemplates(models)
    full = np.asarray(x0 if full_coeffs is None else full_coeffs, dtype=np.float64)
    comp = np.empty(models.nbins)
    composite_(comp, full, models)
    s = comp.sum()
    x0 = np.asarray(x0, dtype=np.float64)
    if s == 0.0:
        return x0.copy()
    return x0 * (_data_sum(data) / s)


def construct_x0(log_age, T_max, normalize_value=1.0):
    """construct_x0 (src/fitting/utilities.jl:31-47): constant-SFR initial
    masses with log_age treated as left bin edges and T_max (Gyr) the last
    right edge."""
    log_age = np.asarray(log_age, dtype=np.float64)
    if not np.log10(T_max) + 9 > log_age.max():
        raise ArgumentError("log10(T_max) + 9 > maximum(logAge)")
    min_la, max_la = log_age.min(), np.log10(T_max) + 9
    uniq = np.concatenate([np.sort(np.unique(log_age)), [max_la]])
    sfr = normalize_value / (10.0 ** max_la - 10.0 ** min_la)
    num = np.array([np.count_nonzero(log_age == la) for la in uniq])
    dt = 10.0 ** uniq[1:] - 10.0 ** uniq[:-1]
    idx = np.searchsorted(uniq, log_age)
    return sfr * dt[idx] / num[idx]


@dataclass
class LogTransformFTResult:
    """LogTransformFTResult (src/fitting/solvers.jl:115-129): coefficients μ,
    standard errors σ from the inverse Hessian in log space, invH, and the
    optimiser's result object."""
    mu: np.ndarray
    sigma: np.ndarray
    invH: np.ndarray
    result: object

    def rand(self, n: int, rng=None):
        """Samples of the Gaussian approximation in log space, exponentiated (:125-129)."""
        rng = np.random.default_rng(rng)
        return np.exp(rng.multivariate_normal(np.log(self.mu), self.invH, size=n).T)


def _bfgs(fun, x0, kws):
    from scipy.optimize import minimize
    opts = {"gtol": 1e-8, "maxiter": 10000}
    opts.update(kws)
    return minimize(fun, x0, jac=True, method="BFGS", options=opts)


def fit_templates(models, data, x0=None, **kws):
    """fit_templates (src/fitting/solvers.jl:172-221): MAP (with the log
    transform's Jacobian) then MLE, both by BFGS in log-coefficients.
    Returns a dict with "map" and "mle" LogTransformFTResults."""
    tpl, hd = _models_data(models, data)
    T = tpl.ntemplates
    x0 = np.ones(T) if x0 is None else np.asarray(x0, dtype=np.float64)
    if x0.size != T:
        raise ArgumentError("axes(coeffs,1) == axes(models,2)")
    x0 = np.log(renormalize_x0(data, tpl, x0))
    G = np.empty(T)
    e, eg = driver_clamps(tpl, data)               # composite::Vector{S} (:173)

    def fg_map(logx):                              # :181-188
        x = np.exp(logx)
        f = fg_(True, G, x, tpl, hd, eps=e, eps_grad=eg) - np.sum(logx)
        return f, G * x - 1.0

    def fg_mle(logx):                              # :189-196
        x = np.exp(logx)
        f = fg_(True, G, x, tpl, hd, eps=e, eps_grad=eg)
        return f, G * x

    r_map = _bfgs(fg_map, x0, kws)
    r_mle = _bfgs(fg_mle, r_map.x, kws)
    out = {}
    for key, r in (("map", r_map), ("mle", r_mle)):
        mu = np.exp(r.x)
        invH = np.asarray(r.hess_inv)
        out[key] = LogTransformFTResult(mu, np.sqrt(np.diag(invH)) * mu, invH, r)
    return out


def fit_templates_fast(models, data, x0=None, **kws):
    """fit_templates_fast (src/fitting/solvers.jl:248-275): MLE with the
    coeffs = θ² transform.  Returns (coeffs, result)."""
    tpl, hd = _models_data(models, data)
    T = tpl.ntemplates
    x0 = np.ones(T) if x0 is None else np.asarray(x0, dtype=np.float64)
    x0 = np.sqrt(renormalize_x0(data, tpl, x0))
    G = np.empty(T)
    e, eg = driver_clamps(tpl, data)               # composite::Vector{S} (:250)

    def fg_mle(sq):                                # :257-264
        x = sq * sq
        f = fg_(True, G, x, tpl, hd, eps=e, eps_grad=eg)
        return f, G * 2.0 * sq

    r = _bfgs(fg_mle, x0, kws)
    return r.x ** 2, r


def fit_templates_lbfgsb(models, data, x0=None, factr=1e-12, pgtol=1e-5, **kws):
    """fit_templates_lbfgsb (src/fitting/solvers.jl:82-90): box-constrained
    L-BFGS-B on the raw coefficients (lb = 0).  Returns (-logL, coeffs)."""
    from scipy.optimize import minimize
    tpl, hd = _models_data(models, data)
    T = tpl.ntemplates
    x0 = np.ones(T) if x0 is None else np.asarray(x0, dtype=np.float64)
    x0 = renormalize_x0(data, tpl, x0)
    G = np.empty(T)
    e, eg = driver_clamps(tpl, data)               # composite::Vector{S} (:84)

    def fg(x):                                     # :88
        f = fg_(True, G, x, tpl, hd, eps=e, eps_grad=eg)
        return f, G.copy()

    opts = {"ftol": factr * np.finfo(np.float64).eps, "gtol": pgtol, "maxiter": 15000, "maxfun": 15000}
    opts.update(kws)
    r = minimize(fg, x0, jac=True, method="L-BFGS-B", bounds=[(0.0, None)] * T, options=opts)
    return r.fun, r.x


class HMCModel:
    """HMCModel + LogDensityProblems interface (src/fitting/hmc_sample.jl:1-37):
    log density in θ = log(coeffs) with the Jacobian term Σθ."""

    def __init__(self, models, data):
        self.models, self.data = _models_data(models, data)
        # hmc_sample allocates composite::Vector{S}, S = eltype(models) (hmc_sample.jl:107)
        self.eps, self.eps_grad = driver_clamps(self.models, data)

    def dimension(self):
        return self.models.ntemplates

    def logdensity(self, theta):                   # :11-21
        theta = np.asarray(theta, dtype=np.float64)
        negll = fg_(True, None, np.exp(theta), self.models, self.data, eps=self.eps, eps_grad=self.eps_grad)
        return -negll + theta.sum()

    def logdensity_and_gradient(self, logx):       # :24-37
        logx = np.asarray(logx, dtype=np.float64)
        x = np.exp(logx)
        G = np.empty_like(x)
        ll = -fg_(True, G, x, self.models, self.data, eps=self.eps, eps_grad=self.eps_grad) + logx.sum()
        return ll, -G * x + 1.0

    def logdensity_and_gradient_batched(self, logX, precision="f64"):
        """The same for every column of logX (T x B) in one batched launch."""
        logX = np.asarray(logX, dtype=np.float64)
        X = np.exp(logX)
        negll, G = fg_batched(X, self.models, self.data, precision=precision, eps=self.eps,
                              eps_grad=self.eps_grad)
        return -negll + logX.sum(axis=0), -G * X + 1.0


class MCMCModel:
    """MCMCModel (src/fitting/mcmc_sample.jl:1-23): log-likelihood of raw
    coefficients, -Inf if any is negative."""

    def __init__(self, models, data):
        self.models, self.data = _models_data(models, data)
        # C::Vector{T}, T = promote_type(eltype(models), eltype(data)) (mcmc_sample.jl:15,20)
        self.eps = driver_clamps(self.models, data)[0]

    def __call__(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        if np.any(theta < 0):
            return -np.inf
        return -fg_(True, None, theta, self.models, self.data, eps=self.eps)

    def batched(self, X, precision="f64"):
        """Log-likelihood of every column of X (T x B); -Inf for columns with a
        negative entry."""
        X = np.asarray(X, dtype=np.float64)
        bad = np.any(X < 0, axis=0)
        Xs = np.where(bad[None, :], 1.0, X)
        negll, _ = fg_batched(Xs, self.models, self.data, precision=precision, want_G=False, eps=self.eps)
        ll = -negll
        ll[bad] = -np.inf
        return ll


def mcmc_sample(models, data, x0, nsteps, a=2.0, rng=None, precision="f64"):
    """Affine-invariant ensemble sampling (the stretch move of
    KissMCMC.emcee used by mcmc_sample, src/fitting/mcmc_sample.jl:97-130),
    with each half-ensemble's proposals evaluated in one batched GPU launch.
    x0: T x nwalkers initial walkers.  Returns (chain nsteps x T x nwalkers,
    log-likelihoods nsteps x nwalkers, acceptance fraction)."""
    model = MCMCModel(models, data)
    rng = np.random.default_rng(rng)
    X = np.array(x0, dtype=np.float64)
    T, W = X.shape
    if W < 4 or W % 2:
        raise ArgumentError("need an even number (>= 4) of walkers")
    lp = model.batched(X, precision)
    chain = np.empty((nsteps, T, W))
    lps = np.empty((nsteps, W))
    acc = 0
    halves = (np.arange(0, W // 2), np.arange(W // 2, W))
    for s in range(nsteps):
        for h in range(2):
            move, other = halves[h], halves[1 - h]
            z = ((a - 1.0) * rng.uniform(size=move.size) + 1.0) ** 2 / a
            partners = X[:, rng.choice(other, size=move.size)]
            prop = partners + z[None, :] * (X[:, move] - partners)
            lpp = model.batched(prop, precision)
            logr = (T - 1) * np.log(z) + lpp - lp[move]
            ok = np.log(rng.uniform(size=move.size)) < logr
            X[:, move[ok]] = prop[:, ok]
            lp[move[ok]] = lpp[ok]
            acc += int(ok.sum())
        chain[s] = X
        lps[s] = lp
    return chain, lps, acc / (nsteps * W)
