"""Hierarchical star-formation-history models on the GPU likelihood.

Mirrors src/fitting/hierarchical/ of StarFormationHistories.jl: metallicity
models (age-metallicity relations LinearAMR / LogarithmicAMR, mass-metallicity
relation PowerLawMZR), the GaussianDispersion of metallicities at fixed age,
calculate_coeffs (per-template coefficients r_{j,k} from per-age stellar
masses R_j), and their fg! -- one GPU fg! (fg_) at the coefficients, then the
chain rule back to (R_j, model parameters) on the host (O(templates)).
``HierarchicalOptimizer.logdensity_and_gradient`` applies the reference's
variable transformations, ``fit_sfh`` runs MAP then MLE, and
``fit_templates_fixed_amr`` fits with fixed per-template relative weights
(fixed_amr.jl).  Vectors are 0-based here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import ArgumentError
from .core import HIPData, HIPTemplates, _eps_of, driver_clamps, fg_

_LN10 = math.log(10.0)


# ------------------------------------------------ metallicity scale (utilities.jl:96-175)

def Y_from_Z(Z, Y_p=0.2485, gamma=1.78):
    """Helium mass fraction Y = Y_p + γ Z (PARSEC)."""
    return Y_p + gamma * Z


def X_from_Z(Z, Y_p=0.2485, gamma=1.78):
    """Hydrogen mass fraction X = 1 - (Y + Z)."""
    return 1 - (Y_from_Z(Z, Y_p, gamma) + Z)


def MH_from_Z(Z, solZ=0.01524, Y_p=0.2485, gamma=1.78):
    """[M/H] = log10(Z/X) - log10(Z/X)_sun (NaN where X <= 0), src/utilities.jl:138-147."""
    X = X_from_Z(Z, Y_p, gamma)
    return math.log10(Z / (X * solZ) * X_from_Z(solZ, Y_p, gamma)) if X > 0 else math.nan


def dMH_dZ(Z, solZ=0.01524, Y_p=0.2485, gamma=1.78):
    """d[M/H]/dZ, src/utilities.jl:152-155."""
    return (Y_p - 1) / (_LN10 * Z * (Y_p + Z + gamma * Z - 1))


def Z_from_MH(MH, solZ=0.01524, Y_p=0.2485, gamma=1.78):
    """Inverse of MH_from_Z: Z = (1 - Y_p) A / (1 + (1 + γ) A), A = 10^([M/H] + log10(Z/X)_sun)."""
    zoverx = 10.0 ** (MH + math.log10(solZ / X_from_Z(solZ, Y_p, gamma)))
    return (1 - Y_p) * zoverx / (1 + (1 + gamma) * zoverx)


# ------------------------------------------------------------ dispersion model

class GaussianDispersion:
    """GaussianDispersion(σ) (dispersion_models.jl:51-110):
    A_{j,k} = exp(-((x_k - μ_j)/σ)^2 / 2)."""

    def __init__(self, sigma, free=(True,)):
        if not sigma > 0:
            raise ArgumentError("σ must be > 0")
        self.sigma = float(sigma)
        self.free = tuple(free)

    nparams = 1

    def fittable_params(self):
        return (self.sigma,)

    def __call__(self, x, mu):
        return np.exp(-((np.asarray(x, dtype=np.float64) - mu) / self.sigma) ** 2 / 2)

    def gradient(self, x, mu):
        """(∂A/∂σ, ∂A/∂μ), arrays over x."""
        A = self(x, mu)
        d = np.asarray(x, dtype=np.float64) - mu
        return np.stack([A * d ** 2 / self.sigma ** 3, A * d / self.sigma ** 2])

    def update_params(self, p):
        return GaussianDispersion(p[0], self.free)

    def transforms(self):
        return (1,)

    def free_params(self):
        return self.free


# ------------------------------------------------------- metallicity models

class AbstractAMR:
    """Age-metallicity relations: μ_j from log10(age) (amr.jl)."""
    nparams = 2


class LinearAMR(AbstractAMR):
    """LinearAMR(α, β, T_max=13.7) (amr.jl:176-247): μ = β + α (T_max - t_Gyr)."""

    def __init__(self, alpha, beta, T_max=13.7, free=(True, True)):
        if alpha < 0:
            raise ArgumentError("α must be ≥ 0")
        if T_max <= 0:
            raise ArgumentError("T_max must be > 0")
        self.alpha, self.beta, self.T_max, self.free = float(alpha), float(beta), float(T_max), tuple(free)

    @classmethod
    def from_constraints(cls, c1, c2, T_max=13.7, free=(True, True)):
        """From ([M/H], lookback Gyr) pairs at two times (amr.jl:224-247)."""
        times, mhs = (c1[1], c2[1]), (c1[0], c2[0])
        dt = times[1] - times[0]
        if dt == 0:
            raise ArgumentError("Constraints are given at identical times.")
        dmh = mhs[1] - mhs[0]
        if not np.sign(dt) != np.sign(dmh):
            raise ArgumentError("The constraints indicate metallicity decreasing towards present-day.")
        alpha = dmh / -dt
        return cls(alpha, min(mhs) - alpha * (T_max - max(times)), T_max, free)

    def fittable_params(self):
        return (self.alpha, self.beta)

    def __call__(self, logAge):
        return self.beta + self.alpha * (self.T_max - 10.0 ** (np.asarray(logAge, dtype=np.float64) - 9))

    def gradient(self, logAge):
        """(∂μ/∂α, ∂μ/∂β)."""
        la = np.asarray(logAge, dtype=np.float64)
        return np.stack([self.T_max - 10.0 ** (la - 9), np.ones_like(la)])

    def update_params(self, p):
        return LinearAMR(p[0], p[1], self.T_max, self.free)

    def transforms(self):
        return (1, 0)

    def free_params(self):
        return self.free


class LogarithmicAMR(AbstractAMR):
    """LogarithmicAMR(α, β, T_max=13.7, MH_func, dMH_dZ) (amr.jl:251-337):
    Z = β + α (T_max - t_Gyr), μ = MH_func(Z)."""

    def __init__(self, alpha, beta, T_max=13.7, MH_func=MH_from_Z, dMH_dZ_func=dMH_dZ, free=(True, True)):
        if alpha < 0:
            raise ArgumentError("α must be ≥ 0")
        if beta < 0:
            raise ArgumentError("β must be ≥ 0")
        if T_max <= 0:
            raise ArgumentError("T_max must be > 0")
        self.alpha, self.beta, self.T_max = float(alpha), float(beta), float(T_max)
        self.MH_func, self.dMH_dZ, self.free = MH_func, dMH_dZ_func, tuple(free)

    @classmethod
    def from_constraints(cls, c1, c2, T_max=13.7, MH_func=MH_from_Z, dMH_dZ_func=dMH_dZ, Z_func=Z_from_MH,
                         free=(True, True)):
        times = (c1[1], c2[1])
        zs = (Z_func(c1[0]), Z_func(c2[0]))
        dt = times[1] - times[0]
        if dt == 0:
            raise ArgumentError("Constraints are given at identical times.")
        dz = zs[1] - zs[0]
        if not np.sign(dt) != np.sign(dz):
            raise ArgumentError("The constraints indicate metallicity decreasing towards present-day.")
        alpha = dz / -dt
        beta = min(zs) - alpha * (T_max - max(times))
        if beta < 0:
            raise ArgumentError("Given constraints result in Z < 0 at T_max.")
        return cls(alpha, beta, T_max, MH_func, dMH_dZ_func, free)

    def fittable_params(self):
        return (self.alpha, self.beta)

    def _z(self, logAge):
        return self.beta + self.alpha * (self.T_max - 10.0 ** (np.asarray(logAge, dtype=np.float64) - 9))

    def __call__(self, logAge):
        z = self._z(logAge)
        return np.vectorize(self.MH_func, otypes=[float])(z)

    def gradient(self, logAge):
        la = np.asarray(logAge, dtype=np.float64)
        d = np.vectorize(self.dMH_dZ, otypes=[float])(self._z(la))
        return np.stack([d * (self.T_max - 10.0 ** (la - 9)), d])

    def update_params(self, p):
        return LogarithmicAMR(p[0], p[1], self.T_max, self.MH_func, self.dMH_dZ, self.free)

    def transforms(self):
        return (1, 1)

    def free_params(self):
        return self.free


class AbstractMZR:
    """Mass-metallicity relations: μ_j from the cumulative stellar mass (mzr.jl)."""
    nparams = 2


class PowerLawMZR(AbstractMZR):
    """PowerLawMZR(α, MH0, logMstar0=6) (mzr.jl:217-285):
    μ = MH0 + α (log10(M*) - logMstar0)."""

    def __init__(self, alpha, MH0, logMstar0=6.0, free=(True, True)):
        if alpha < 0:
            raise ArgumentError("α must be ≥ 0")
        self.alpha, self.MH0, self.logMstar0, self.free = float(alpha), float(MH0), float(logMstar0), tuple(free)

    def fittable_params(self):
        return (self.alpha, self.MH0)

    def __call__(self, Mstar):
        return self.MH0 + self.alpha * (np.log10(np.asarray(Mstar, dtype=np.float64)) - self.logMstar0)

    def gradient(self, Mstar):
        """(∂μ/∂α, ∂μ/∂MH0, ∂μ/∂M*)."""
        M = np.asarray(Mstar, dtype=np.float64)
        return np.stack([np.log10(M) - self.logMstar0, np.ones_like(M), self.alpha / M / _LN10])

    def update_params(self, p):
        return PowerLawMZR(p[0], p[1], self.logMstar0, self.free)

    def transforms(self):
        return (1, 0)

    def free_params(self):
        return self.free


# ------------------------------------------------------------ coefficients

def _groups(logAge):
    """unique(logAge) in first-appearance order, the index of each entry's age, and the
    entries of each age."""
    logAge = np.asarray(logAge, dtype=np.float64)
    _, first, inv = np.unique(logAge, return_index=True, return_inverse=True)
    order = np.argsort(first)                 # unique() keeps first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    jidx = rank[inv]
    return logAge[np.sort(first)], jidx


def _cum_mstar(mstars, ula):
    """Cumulative stellar mass from the oldest age bin (mzr.jl:71-75)."""
    s = np.argsort(-ula, kind="stable")
    cm = np.empty_like(mstars)
    cm[s] = np.cumsum(mstars[s])
    return cm, s


def calculate_coeffs(MH_model, disp_model, variables, logAge, metallicities):
    """calculate_coeffs (amr.jl:56-77, mzr.jl:58-84): r_{j,k} = R_j A_{j,k} / sum_k A_{j,k}."""
    variables = np.asarray(variables, dtype=np.float64)
    mh = np.asarray(metallicities, dtype=np.float64)
    ula, jidx = _groups(logAge)
    if variables.size != ula.size:
        raise ArgumentError("Length of `variables` must be the same as `unique_logAge`.")
    if mh.size != jidx.size:
        raise ArgumentError("length(logAge) == length(metallicities)")
    if isinstance(MH_model, AbstractMZR):
        mu = MH_model(_cum_mstar(variables, ula)[0])
    else:
        mu = MH_model(ula)
    A = disp_model(mh, mu[jidx])
    Asum = np.bincount(jidx, weights=A, minlength=ula.size)
    return A * variables[jidx] / Asum[jidx]


# ------------------------------------------------------------------- fg!

def fg_hierarchical(F, G, MH_model0, disp_model0, variables, models: HIPTemplates, data, logAge, metallicities):
    """fg! of a hierarchical model (see _fg_hierarchical)."""
    # trial points of a line search can underflow every A_{j,k} of an age
    # (tiny σ); the result is then non-finite, as in the reference
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return _fg_hierarchical(F, G, MH_model0, disp_model0, variables, models, data, logAge, metallicities)


def _fg_hierarchical(F, G, MH_model0, disp_model0, variables, models, data, logAge, metallicities):
    """fg!(F, G, MH_model0, disp_model0, variables, models, data, composite,
    logAge, metallicities) -- amr.jl:82-171 / mzr.jl:88-214: -logL at the
    coefficients of `variables` = (R_j..., MH-model params..., dispersion
    params...) and, if G is given, its gradient written into G.  The
    likelihood and ∇ over coefficients come from one GPU fg!."""
    variables = np.asarray(variables, dtype=np.float64)
    zpar, dpar = MH_model0.nparams, disp_model0.nparams
    nR = variables.size - zpar - dpar
    MH = MH_model0.update_params(variables[nR:nR + zpar])
    disp = disp_model0.update_params(variables[nR + zpar:])
    R = variables[:nR]
    mh = np.asarray(metallicities, dtype=np.float64)
    ula, jidx = _groups(logAge)
    if nR != ula.size:
        raise ArgumentError("Length of `variables` must be the same as `unique_logAge`.")
    mzr = isinstance(MH_model0, AbstractMZR)
    if mzr:
        cum, s_idxs = _cum_mstar(R, ula)
        mu = MH(cum)
        gradmu = MH.gradient(cum)                   # (zpar + 1, nR): params..., ∂μ/∂M*
    else:
        mu = MH(ula)
        gradmu = MH.gradient(ula)                   # (zpar, nR)
    A = disp(mh, mu[jidx])
    Asum = np.bincount(jidx, weights=A, minlength=nR)
    coeffs = A * R[jidx] / Asum[jidx]
    Gc = np.empty(coeffs.size) if G is not None else None
    # composite = similar(data, S), S = promote_type(eltype(models), eltype(data))
    # (generic_fitting.jl:97-98): both clamps are eps(S)
    e = _eps_of(models, data)
    nll = fg_(True, Gc, coeffs, models, data, eps=e, eps_grad=e)
    if G is None:
        return nll if F is not None else None
    if G.size != variables.size:
        raise ArgumentError("axes(G) == axes(variables)")
    fullG = -Gc                                      # ∇logL over coefficients
    gd = disp.gradient(mh, mu[jidx])                 # (dpar + 1, ntpl): params..., ∂A/∂μ
    dA_dmu = gd[-1]
    ks = lambda v: np.bincount(jidx, weights=v, minlength=nR)
    G[:] = 0.0
    if not mzr:
        # ∂F/∂R_j = -sum_k fullG_k r_k / R_j                                      (amr.jl:132)
        G[:nR] = -ks(fullG * coeffs) / R
        ksum_dA = ks(dA_dmu)
        psum = -ks(fullG * R[jidx] / Asum[jidx] * (dA_dmu - A / Asum[jidx] * ksum_dA[jidx]))   # :145-146
        for p in range(zpar):
            if MH.free_params()[p]:
                G[nR + p] += np.sum(psum * gradmu[p])
    else:
        dmu_dR = gradmu[-1]
        dA_dR = dA_dmu * dmu_dR[jidx]                                          # mzr.jl:159
        ksum_dA_dR = ks(dA_dR)
        drjk = fullG * (R[jidx] / Asum[jidx] * (dA_dR - A * ksum_dA_dR[jidx] / Asum[jidx]))    # :163-164
        ksum_drjk = ks(drjk)
        cum_d = np.cumsum(ksum_drjk[s_idxs][::-1])[::-1]                         # :169
        for i in range(1, nR):                                                   # :175-177
            G[s_idxs[i - 1]] -= cum_d[i]
        G[:nR] -= ks(fullG * (coeffs / R[jidx] + (dA_dR - ksum_dA_dR[jidx] * A / Asum[jidx]) * R[jidx] / Asum[jidx]))  # :184-186
        ksum_dA = ks(dA_dmu)
        psum = -ks(fullG * R[jidx] / Asum[jidx] * (dA_dmu - A / Asum[jidx] * ksum_dA[jidx]))   # :189-191
        for p in range(zpar):
            if MH.free_params()[p]:
                G[nR + p] += np.sum(psum * gradmu[p])
    for p in range(dpar):                                                        # amr.jl:152-161
        if disp.free_params()[p]:
            dP = gd[p]
            ksum_dP = ks(dP)
            G[nR + zpar + p] -= np.sum(ks(fullG * R[jidx] / Asum[jidx] * (dP - A / Asum[jidx] * ksum_dP[jidx])))
    return nll if F is not None else None


# ------------------------------------------------------------ log density

@dataclass
class HierarchicalOptimizer:
    """HierarchicalOptimizer (generic_fitting.jl:43-60) + its
    LogDensityProblems interface: variables are log(R_j), then the free model
    parameters, log-transformed where transforms() says 1."""
    MH_model0: object
    disp_model0: object
    models: HIPTemplates
    data: object
    logAge: np.ndarray
    metallicities: np.ndarray
    jacobian_corrections: bool = True

    def _layout(self):
        tf = np.array(self.MH_model0.transforms() + self.disp_model0.transforms())
        free = np.array(self.MH_model0.free_params() + self.disp_model0.free_params(), dtype=bool)
        init = np.array(self.MH_model0.fittable_params() + self.disp_model0.fittable_params(), dtype=np.float64)
        return tf, free, init

    def dimension(self):
        tf, free, _ = self._layout()
        return np.unique(self.logAge).size + int(free.sum())

    def logdensity_and_gradient(self, xvec):
        """(logL [+ Jacobian terms], ∇ wrt xvec) -- generic_fitting.jl:100-199."""
        xvec = np.asarray(xvec, dtype=np.float64)
        tf, free, init = self._layout()
        npar = tf.size
        nR = xvec.size - int(free.sum())
        x = np.empty(nR + npar)
        x[:nR] = np.exp(xvec[:nR])
        par = xvec[nR:]
        tfree = tf[free]
        xz = np.where(tfree == 1, np.exp(par), np.where(tfree == -1, -np.exp(par), par))
        x[nR:][free] = xz
        x[nR:][~free] = init[~free]
        G2 = np.empty_like(x)
        nll = fg_hierarchical(True, G2, self.MH_model0, self.disp_model0, x, self.models, self.data,
                              self.logAge, self.metallicities)
        pos = np.concatenate([np.arange(nR), nR + np.flatnonzero((tf == 1) & free)])
        neg = nR + np.flatnonzero((tf == -1) & free)
        if self.jacobian_corrections:
            nll -= np.sum(np.log(x[pos]))
            G2[pos] = G2[pos] * x[pos] - 1
            nll += np.sum(np.log(-x[neg]))
            G2[neg] = -G2[neg] * x[neg] + 1
        else:
            G2[pos] = G2[pos] * x[pos]
            G2[neg] = -G2[neg] * x[neg]
        G = np.concatenate([G2[:nR], G2[nR:][free]])
        return -nll, -G

    def logdensity(self, xvec):
        return self.logdensity_and_gradient(xvec)[0]


def _to_transformed(MH_model0, disp_model0, R0):
    tf = np.array(MH_model0.transforms() + disp_model0.transforms())
    free = np.array(MH_model0.free_params() + disp_model0.free_params(), dtype=bool)
    init = np.array(MH_model0.fittable_params() + disp_model0.fittable_params(), dtype=np.float64)
    p = np.where(tf == 1, np.log(np.abs(init)), np.where(tf == -1, np.log(np.abs(init)), init))[free]
    return np.concatenate([np.log(np.asarray(R0, dtype=np.float64)), p]), tf, free, init


@dataclass
class HierarchicalResult:
    """One stage of fit_sfh: the fitted R_j, MH-model and dispersion
    parameters (all, fixed ones included), σ from the BFGS inverse Hessian in
    the transformed variables, and the optimiser's result."""
    mstars: np.ndarray
    MH_params: np.ndarray
    disp_params: np.ndarray
    invH: np.ndarray
    sigma: np.ndarray
    result: object


def fit_sfh(MH_model0, disp_model0, models, data, logAge, metallicities, x0=None, gtol=1e-8, **kws):
    """fit_sfh (generic_fitting.jl:251-334): BFGS on the transformed
    variables, MAP (Jacobian corrections) then MLE from the MAP optimum.
    Returns {"map": HierarchicalResult, "mle": ...}."""
    from scipy.optimize import minimize
    tpl = models if isinstance(models, HIPTemplates) else HIPTemplates(models)
    hd = data if isinstance(data, HIPData) else HIPData(np.asarray(data, dtype=np.float64).ravel(order="F"),
                                                         ctx=tpl.ctx)
    logAge = np.asarray(logAge, dtype=np.float64)
    metallicities = np.asarray(metallicities, dtype=np.float64)
    ula, _ = _groups(logAge)
    if x0 is None:
        from .fitting import construct_x0
        x0 = construct_x0(ula, 13.7, normalize_value=1e6)
    xt, tf, free, init = _to_transformed(MH_model0, disp_model0, x0)
    out = {}
    for key, jac in (("map", True), ("mle", False)):
        prob = HierarchicalOptimizer(MH_model0, disp_model0, tpl, hd, logAge, metallicities, jac)

        def fun(v):
            ll, g = prob.logdensity_and_gradient(v)
            return -ll, -g

        opts = {"gtol": gtol, "maxiter": 10000, "norm": np.inf}
        opts.update(kws)
        r = minimize(fun, xt, jac=True, method="BFGS", options=opts)
        xt = r.x
        nR = ula.size
        full = init.copy()
        full[free] = np.where(tf[free] == 1, np.exp(r.x[nR:]), np.where(tf[free] == -1, -np.exp(r.x[nR:]), r.x[nR:]))
        zpar = MH_model0.nparams
        invH = np.asarray(r.hess_inv)
        out[key] = HierarchicalResult(np.exp(r.x[:nR]), full[:zpar], full[zpar:], invH, np.sqrt(np.diag(invH)), r)
    return out


def fit_templates_fixed_amr(models, data, logAge, metallicities, relweights, x0=None, gtol=1e-8, **kws):
    """fixed_amr (fixed_amr.jl:64-155): per-age stellar masses with fixed
    per-template relative weights (normalised per age, :82-93), MAP then MLE
    in log space.  Returns {"map": (mu, sigma, result), "mle": ...}."""
    from scipy.optimize import minimize
    tpl = models if isinstance(models, HIPTemplates) else HIPTemplates(models)
    hd = data if isinstance(data, HIPData) else HIPData(np.asarray(data, dtype=np.float64).ravel(order="F"),
                                                         ctx=tpl.ctx)
    ula, jidx = _groups(logAge)
    rw = np.asarray(relweights, dtype=np.float64)
    if rw.size != jidx.size or tpl.ntemplates != jidx.size:
        raise ArgumentError("length(logAge) == length(metallicities) == length(relweights) == ntemplates")
    rw = rw / np.bincount(jidx, weights=rw, minlength=ula.size)[jidx]        # :68-79
    if x0 is None:
        from .fitting import construct_x0
        x0 = construct_x0(ula, 13.7, normalize_value=1e6)
    x0 = np.asarray(x0, dtype=np.float64)
    if not isinstance(data, HIPData):
        # renormalize_x0 with the full per-template coefficients (:85-91)
        from .fitting import renormalize_x0
        x0 = renormalize_x0(np.asarray(data, dtype=np.float64), tpl, x0, rw * x0[jidx])
    Gc = np.empty(jidx.size)
    e, eg = driver_clamps(tpl, data)               # composite::Vector{S} (fixed_amr.jl:50)

    def fg(xvec, mapj):
        x = np.exp(xvec)
        coeffs = rw * x[jidx]
        nll = fg_(True, Gc, coeffs, tpl, hd, eps=e, eps_grad=eg)
        g = np.bincount(jidx, weights=Gc * coeffs, minlength=ula.size)          # :122, :145
        return (nll - np.sum(xvec), g - 1.0) if mapj else (nll, g)

    opts = {"gtol": gtol, "maxiter": 10000, "norm": np.inf}
    opts.update(kws)
    r1 = minimize(fg, np.log(x0), jac=True, method="BFGS", args=(True,), options=opts)
    r2 = minimize(fg, r1.x, jac=True, method="BFGS", args=(False,), options=opts)
    return {k: (np.exp(r.x), np.sqrt(np.diag(np.asarray(r.hess_inv))), r) for k, r in (("map", r1), ("mle", r2))}
