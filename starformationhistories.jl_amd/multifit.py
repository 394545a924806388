"""Many independent fits of uneven size, evaluated and optimised together on
the GPU (sfh_multifit_* of libsfh_hip.so).

``MultiFit(models, data)`` groups fits (fit f: templates ``models[f]`` of
shape (nbins_f, ntemplates_f), observed counts ``data[f]``).  Its ``fg``
evaluates fg! (src/fitting/solvers.jl:20-38) of every fit in one launch, and
``fit_templates`` / ``fit_templates_fast`` run the optimisation of
src/fitting/solvers.jl:172-221 / 248-275 for every fit at once, the
quasi-Newton iterations on the device in captured HIP graphs.  Results come
back per fit in the reference's shapes (``LogTransformFTResult`` with μ, σ,
invH for the MAP and MLE stages).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError
from .core import Context, HIPData, HIPTemplates, _f64eps, default_context
from .fitting import LogTransformFTResult


@dataclass
class StageInfo:
    """What the optimiser reports for one stage of one fit."""
    status: str
    iterations: int
    evaluations: int
    minimum: float


class MultiFit:
    """A group of independent fits on one context.  ``models`` / ``data``:
    sequences of HIPTemplates / HIPData handles or of host arrays (uploaded
    here).  Each fit may have its own template and bin counts (at most 1024
    templates per fit)."""

    def __init__(self, models, data, *, ctx: Context | None = None):
        if len(models) != len(data) or len(models) == 0:
            raise ArgumentError("models and data must be non-empty sequences of equal length")
        self.ctx = ctx or (models[0].ctx if isinstance(models[0], HIPTemplates) else default_context())
        self.templates = [m if isinstance(m, HIPTemplates) else HIPTemplates(m, ctx=self.ctx) for m in models]
        self.data = [d if isinstance(d, HIPData) else HIPData(d, ctx=self.ctx) for d in data]
        self.sizes = np.array([t.ntemplates for t in self.templates], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.hoffsets = np.concatenate([[0], np.cumsum(self.sizes ** 2)])
        tp = (C.c_void_p * len(models))(*[t.handle.value for t in self.templates])
        dp = (C.c_void_p * len(models))(*[d.handle.value for d in self.data])
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.sfh_multifit_create(self.ctx.handle, len(models), tp, dp, C.byref(h)))
        self.handle = h

    @property
    def nfits(self) -> int:
        return len(self.templates)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.sfh_multifit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ shapes
    def shape(self):
        """(item_elems, waves, wg_per_cu) of the grouped fg! launch."""
        ie, nw, wg = C.c_int64(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.lib.sfh_multifit_get_shape(self.handle, C.byref(ie), C.byref(nw), C.byref(wg)))
        return ie.value, nw.value, wg.value

    def set_shape(self, item_elems: int, waves: int, wg_per_cu: int):
        self.ctx.check(self.ctx.lib.sfh_multifit_set_shape(self.handle, int(item_elems), int(waves), int(wg_per_cu)))

    def autotune(self, reps: int = 3):
        """Time the candidate launch shapes on these fits and keep the fastest."""
        ie, nw, wg = C.c_int64(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.lib.sfh_multifit_autotune(self.handle, int(reps), C.byref(ie), C.byref(nw),
                                                          C.byref(wg)))
        return ie.value, nw.value, wg.value

    # ------------------------------------------------------------ vectors
    def split(self, v):
        """Split a concatenated per-template vector into one array per fit."""
        return [v[self.offsets[f]:self.offsets[f + 1]] for f in range(self.nfits)]

    def concat(self, vectors, name="coefficients"):
        if len(vectors) != self.nfits:
            raise ArgumentError(f"need one {name} vector per fit")
        out = np.empty(int(self.offsets[-1]))
        for f, v in enumerate(vectors):
            v = np.asarray(v, dtype=np.float64).ravel()
            if v.size != self.sizes[f]:
                raise ArgumentError(f"axes({name},1) == axes(models,2) fails for fit {f}")
            out[self.offsets[f]:self.offsets[f + 1]] = v
        return out

    # ------------------------------------------------------------ fg!
    def fg(self, coeffs, want_G: bool = True, eps: float = _f64eps, eps_grad: float | None = None):
        """fg!(true, G, coeffs_f, models_f, data_f) for every fit: returns
        (-logL per fit, list of -∇logL per fit or None).  ``eps`` clamps the
        log-likelihood, ``eps_grad`` (default ``eps``) the gradient, as sfh_fg."""
        eps_grad = eps if eps_grad is None else eps_grad
        X = self.concat(coeffs)
        nl = np.empty(self.nfits)
        G = np.empty(X.size) if want_G else None
        self.ctx.check(self.ctx.lib.sfh_multifit_fg(self.handle, X.ctypes.data, float(eps), float(eps_grad),
                                                    nl.ctypes.data,
                                                    G.ctypes.data if want_G else None))
        return nl, (self.split(G) if want_G else None)

    def fg_device(self, d_X, d_negLL, d_G=None, eps: float = _f64eps, eps_grad: float | None = None):
        """Asynchronous form on device buffers (concatenated X and G)."""
        eps_grad = eps if eps_grad is None else eps_grad
        from .core import _ptr
        self.ctx.check(self.ctx.lib.sfh_multifit_fg_device(self.handle, _ptr(d_X), float(eps), float(eps_grad),
                                                           _ptr(d_negLL),
                                                           _ptr(d_G)))

    # ------------------------------------------------------------ fits
    def _fit(self, x0, transform, method, memory, g_abstol, iterations, steps_per_graph, eps, want_invH,
             eps_grad=None):
        opt = _lib.sfh_fit_options()
        self.ctx.lib.sfh_fit_options_default(C.byref(opt))
        opt.transform = transform
        opt.method = method
        opt.memory = int(memory)
        opt.g_abstol = float(g_abstol)
        opt.max_iterations = int(iterations)
        opt.steps_per_graph = int(steps_per_graph)
        opt.eps = float(eps)
        opt.eps_grad = float(eps if eps_grad is None else eps_grad)
        nT = int(self.offsets[-1])
        nH = int(self.hoffsets[-1])
        mu_map = np.zeros(nT)
        mu_mle = np.zeros(nT)
        bfgs = method == _lib.SFH_OPT_BFGS
        invH_map = np.zeros(nH) if (want_invH and bfgs and transform == _lib.SFH_FIT_LOG) else None
        invH_mle = np.zeros(nH) if (want_invH and bfgs) else None
        fmin = np.zeros(2 * self.nfits)
        iters = np.zeros(2 * self.nfits, dtype=np.int64)
        evals = np.zeros(2 * self.nfits, dtype=np.int64)
        status = np.zeros(2 * self.nfits, dtype=np.int32)
        res = _lib.sfh_fit_results(mu_map.ctypes.data, mu_mle.ctypes.data,
                                   invH_map.ctypes.data if invH_map is not None else None,
                                   invH_mle.ctypes.data if invH_mle is not None else None,
                                   fmin.ctypes.data, iters.ctypes.data, evals.ctypes.data, status.ctypes.data, 0)
        X0 = self.concat(x0, "x0") if x0 is not None else None
        self.ctx.check(self.ctx.lib.sfh_multifit_fit(self.handle, X0.ctypes.data if X0 is not None else None,
                                                     C.byref(opt), C.byref(res)))
        self.steps = int(res.steps)
        return mu_map, mu_mle, invH_map, invH_mle, fmin, iters, evals, status

    def _info(self, stage, f, fmin, iters, evals, status):
        k = stage * self.nfits + f
        return StageInfo(_lib.FIT_STATUS.get(int(status[k]), str(status[k])), int(iters[k]), int(evals[k]),
                         float(fmin[k]))

    def _invH(self, H, f):
        if H is None:
            return None
        T = int(self.sizes[f])
        return H[self.hoffsets[f]:self.hoffsets[f + 1]].reshape(T, T)

    def fit_templates(self, x0=None, method: str = "bfgs", memory: int = 10, g_abstol: float = 1e-8,
                      iterations: int = 1000, steps_per_graph: int = 16, eps: float = _f64eps,
                      eps_grad: float | None = None):
        """fit_templates (src/fitting/solvers.jl:172-221) for every fit:
        returns a list of {"map": LogTransformFTResult, "mle": ...}.  With
        method="lbfgs" the inverse Hessian (and so σ) is not available.
        ``eps`` / ``eps_grad`` are the log-likelihood's and the gradient's clamps
        (eps(Float64) both by default; the reference's composite has the
        models' element type, solvers.jl:173)."""
        m = {"bfgs": _lib.SFH_OPT_BFGS, "lbfgs": _lib.SFH_OPT_LBFGS}[method]
        mu_map, mu_mle, Hmap, Hmle, fmin, iters, evals, status = self._fit(
            x0, _lib.SFH_FIT_LOG, m, memory, g_abstol, iterations, steps_per_graph, eps, True, eps_grad)
        out = []
        for f in range(self.nfits):
            sl = slice(self.offsets[f], self.offsets[f + 1])
            r = {}
            for key, mu, H, stage in (("map", mu_map, Hmap, 0), ("mle", mu_mle, Hmle, 1)):
                invH = self._invH(H, f)
                sigma = np.sqrt(np.diag(invH)) * mu[sl] if invH is not None else None   # solvers.jl:206-214
                r[key] = LogTransformFTResult(mu[sl].copy(), sigma, invH,
                                              self._info(stage, f, fmin, iters, evals, status))
            out.append(r)
        return out

    def fit_templates_fast(self, x0=None, method: str = "bfgs", memory: int = 10, g_abstol: float = 1e-8,
                           iterations: int = 1000, steps_per_graph: int = 16, eps: float = _f64eps,
                           eps_grad: float | None = None):
        """fit_templates_fast (src/fitting/solvers.jl:248-275) for every fit:
        returns a list of (coeffs, StageInfo) with coeffs = θ² at the MLE."""
        m = {"bfgs": _lib.SFH_OPT_BFGS, "lbfgs": _lib.SFH_OPT_LBFGS}[method]
        _, mu_mle, _, _, fmin, iters, evals, status = self._fit(
            x0, _lib.SFH_FIT_SQRT, m, memory, g_abstol, iterations, steps_per_graph, eps, False, eps_grad)
        return [(mu_mle[self.offsets[f]:self.offsets[f + 1]].copy(), self._info(1, f, fmin, iters, evals, status))
                for f in range(self.nfits)]
