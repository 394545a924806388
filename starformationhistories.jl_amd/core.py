"""Template-mixture likelihood on the GPU, behind the reference's API.

Mirrors src/fitting/fitting_base.jl and src/fitting/solvers.jl of
StarFormationHistories.jl v1.3.1.  Julia names that are not Python
identifiers map as:  composite! -> composite_,  ∇loglikelihood ->
grad_loglikelihood,  ∇loglikelihood! -> grad_loglikelihood_,  fg! -> fg_.

Templates live on the device in a ``HIPTemplates`` handle (the flattened
``models::AbstractMatrix`` layout, nbins x ntemplates, FP32); every numerical
call goes through libsfh_hip.so.  Host numpy arrays are copied per call;
``*_device`` functions take device pointers (e.g. torch tensors) and stay
asynchronous on the context's stream.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import ArgumentError, check

_f64eps = float(np.finfo(np.float64).eps)


def _ptr(a) -> int:
    """Host (numpy) or device (torch) data pointer."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, int):
        return a
    raise TypeError(f"cannot take a pointer of {type(a)}")


class Context:
    """One device + one HIP stream (an sfh_ctx).  Calls on one context are
    serialised; give each host thread its own context (``default_context``
    does), as the reference gives each task its own buffers
    (src/fitting/hmc_sample.jl:127)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.sfh_ctx_create(int(device), C.c_void_p(stream) if stream else None, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sfh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        check(rc, self.handle)

    def synchronize(self):
        self.check(self.lib.sfh_ctx_synchronize(self.handle))

    def set_timing(self, on: bool):
        self.check(self.lib.sfh_ctx_set_timing(self.handle, 1 if on else 0))

    def kernel_stats(self, name: str):
        ms = C.c_double()
        n = C.c_int64()
        self.check(self.lib.sfh_ctx_kernel_stats(self.handle, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def valu_issue_rate(self, iters: int = 20000) -> float:
        """VALU lane-instructions per second that the deposition strip's
        instruction mix reaches on this device (four waves per SIMD)."""
        r = C.c_double()
        self.check(self.lib.sfh_probe_valu_issue(self.handle, int(iters), C.byref(r)))
        return r.value

    def reset_stats(self):
        self.check(self.lib.sfh_ctx_reset_stats(self.handle))


_tls = threading.local()


def default_context() -> Context:
    """Per-thread context on device 0."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


def stack_models(models) -> np.ndarray:
    """stack_models (src/fitting/utilities.jl:12-13): vec each template
    (column-major) and hcat -> nbins x ntemplates, Fortran order."""
    cols = [np.asarray(m).ravel(order="F") for m in models]
    return np.asfortranarray(np.stack(cols, axis=1))


class HIPTemplates:
    """Device-resident templates (nbins x ntemplates FP32), the GPU stand-in
    for the flattened ``models::AbstractMatrix``.  Accepts an (N, T) array,
    a list of 2-D templates (stacked like stack_models) or a device pointer
    with explicit shape."""

    def __init__(self, models=None, *, ctx: Context | None = None, device_ptr: int | None = None,
                 shape: tuple | None = None):
        self.ctx = ctx or default_context()
        lib = self.ctx.lib
        h = C.c_void_p()
        if device_ptr is not None:
            nbins, ntpl = shape
            self.ctx.check(lib.sfh_templates_create_device(self.ctx.handle, C.c_void_p(device_ptr),
                                                           int(nbins), int(ntpl), C.byref(h)))
        else:
            if isinstance(models, (list, tuple)):
                models = stack_models(models)
            src_dtype = np.asarray(models).dtype
            m = np.asfortranarray(models, dtype=np.float32)
            if m.ndim != 2:
                raise ArgumentError("models must be a 2-D (nbins, ntemplates) array")
            nbins, ntpl = m.shape
            self.ctx.check(lib.sfh_templates_create(self.ctx.handle, m.ctypes.data, nbins, ntpl,
                                                    C.byref(h)))
        self.handle = h
        self.shape = (int(nbins), int(ntpl))
        # The element type of the caller's models: the reference's drivers
        # allocate their composite with it (fit_templates, solvers.jl:173;
        # hmc_sample, hmc_sample.jl:107), which fixes the gradient's clamp
        # eps(eltype(composite)).  Device-pointer templates are FP32.
        self.eltype = np.dtype(np.float32)
        if device_ptr is None:
            self.eltype = np.dtype(np.float32) if src_dtype == np.float32 else np.dtype(np.float64)

    @property
    def nbins(self):
        return self.shape[0]

    @property
    def ntemplates(self):
        return self.shape[1]

    def device_ptr(self) -> int:
        """Device address of the resident FP32 models (ld x ntemplates,
        ld = nbins rounded up to 256, zero padded)."""
        return self.ctx.lib.sfh_templates_device_ptr(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.sfh_templates_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HIPData:
    """Device-resident observed Hess diagram (FP64 counts)."""

    def __init__(self, data=None, *, ctx: Context | None = None, device_ptr: int | None = None,
                 nbins: int | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        if device_ptr is not None:
            self.ctx.check(self.ctx.lib.sfh_data_create_device(self.ctx.handle, C.c_void_p(device_ptr),
                                                               int(nbins), C.byref(h)))
            n = int(nbins)
        else:
            d = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel(order="F"))
            n = d.size
            self.ctx.check(self.ctx.lib.sfh_data_create(self.ctx.handle, d.ctypes.data, n, C.byref(h)))
        self.handle = h
        self.nbins = n

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.sfh_data_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_data(data, models: HIPTemplates) -> HIPData:
    if isinstance(data, HIPData):
        return data
    return HIPData(data, ctx=models.ctx)


def _vec64(a, n=None, name="array"):
    v = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
    if n is not None and v.size != n:
        raise ArgumentError(f"length of {name} ({v.size}) != {n}")
    return v


def _dtype_of(a):
    """Element type of a host array, a dtype, or a handle (HIPTemplates: the
    caller's models' type; HIPData: FP64 counts)."""
    if isinstance(a, HIPTemplates):
        return a.eltype
    if isinstance(a, HIPData):
        return np.dtype(np.float64)
    if isinstance(a, np.dtype):
        return a
    return np.asarray(a).dtype


def _eps_of(*arrays) -> float:
    """eps(promote_type(...)) of the reference's clamps (fitting_base.jl:86,148):
    eps(Float32) when every element type is Float32 or an integer (with at
    least one Float32), else eps(Float64)."""
    kinds = [_dtype_of(a) for a in arrays if a is not None]
    if kinds and all(k in (np.float32, np.float16) or np.issubdtype(k, np.integer) for k in kinds) \
            and any(k == np.float32 for k in kinds):
        return float(np.finfo(np.float32).eps)
    return _f64eps


def driver_clamps(models: "HIPTemplates", data) -> tuple:
    """(eps, eps_grad) of fg! as the reference's drivers call it: their
    composite is a Vector of the models' element type S (fit_templates,
    solvers.jl:173; hmc_sample, hmc_sample.jl:107), so the log-likelihood
    clamps with eps(promote_type(S, eltype(data))) (fitting_base.jl:86) and the
    gradient with eps(S) (:277)."""
    return _eps_of(models, data), _eps_of(models.eltype)


def composite_(composite: np.ndarray, coeffs, models: HIPTemplates):
    """composite!(composite, coeffs, models) -- src/fitting/fitting_base.jl:55-65.
    Overwrites ``composite`` (float64, length nbins) with models * coeffs."""
    if not isinstance(models, HIPTemplates):
        raise TypeError("models must be a HIPTemplates (device-resident) matrix")
    N, T = models.shape
    c = _vec64(coeffs, T, "coeffs")
    if composite.shape != (N,) and composite.size != N:
        raise ArgumentError("axes(composite,1) == axes(models,1)")
    out = np.empty(N, dtype=np.float64)
    models.ctx.check(models.ctx.lib.sfh_composite(models.ctx.handle, models.handle, c.ctypes.data,
                                                  out.ctypes.data))
    composite[...] = out.reshape(composite.shape, order="F")
    return None


def loglikelihood(*args, eps: float | None = None):
    """loglikelihood(composite, data) -- src/fitting/fitting_base.jl:84-96, or
    loglikelihood(coeffs, models::HIPTemplates, data) -- :116-125."""
    if len(args) == 2:
        composite, data = args
        cm = np.asarray(composite)
        dt = np.asarray(data)
        if cm.shape != dt.shape:
            raise ArgumentError("axes(composite) == axes(data)")
        e = _eps_of(cm, dt) if eps is None else eps
        c64, d64 = _vec64(cm), _vec64(dt)
        out = C.c_double()
        ctx = default_context()
        ctx.check(ctx.lib.sfh_loglikelihood(ctx.handle, c64.ctypes.data, d64.ctypes.data, c64.size, e,
                                            C.byref(out)))
        return out.value
    if len(args) == 3:
        coeffs, models, data = args
        if not isinstance(models, HIPTemplates):
            raise TypeError("models must be a HIPTemplates")
        d = _as_data(data, models)
        if d.nbins != models.nbins:
            raise ArgumentError("axes(data,1) == axes(models,1)")
        # composite = Vector{S}, S = promote_type(coeffs, models, data) (:121)
        e = _eps_of(coeffs, models, data) if eps is None else eps
        c = _vec64(coeffs, models.ntemplates, "coeffs")
        out = C.c_double()
        models.ctx.check(models.ctx.lib.sfh_loglikelihood_coeffs(models.ctx.handle, models.handle, d.handle,
                                                                 c.ctypes.data, e, C.byref(out)))
        return out.value
    raise TypeError("loglikelihood takes (composite, data) or (coeffs, models, data)")


def grad_loglikelihood(*args, eps: float | None = None) -> np.ndarray:
    """∇loglikelihood(models, composite, data) -- src/fitting/fitting_base.jl:177-182,
    or ∇loglikelihood(coeffs, models, data) -- :202-211."""
    if len(args) != 3:
        raise TypeError("grad_loglikelihood takes 3 arguments")
    a, b, c = args
    if isinstance(a, HIPTemplates):
        models, composite, data = a, b, c
        comp_dtype = _dtype_of(composite)
        comp = _vec64(composite, models.nbins, "composite")
    elif isinstance(b, HIPTemplates):
        coeffs, models, data = a, b, c
        # composite = Vector{S}, S = promote_type(coeffs, models, data) (:207)
        comp_dtype = np.result_type(_dtype_of(coeffs), models.eltype, _dtype_of(data))
        comp = np.empty(models.nbins, dtype=np.float64)
        composite_(comp, coeffs, models)
    else:
        raise TypeError("one argument must be a HIPTemplates")
    # eps(promote_type(eltype(model), eltype(composite), eltype(data))) (:148)
    e = _eps_of(models, comp_dtype, data) if eps is None else eps
    d = _as_data(data, models)
    if d.nbins != models.nbins:
        raise ArgumentError("axes(composite,1) == axes(data,1) == axes(models,1)")
    G = np.empty(models.ntemplates, dtype=np.float64)
    models.ctx.check(models.ctx.lib.sfh_grad_loglikelihood(models.ctx.handle, models.handle, comp.ctypes.data,
                                                           d.handle, e, G.ctypes.data))
    return G


def grad_loglikelihood_(G: np.ndarray, composite: np.ndarray, models: HIPTemplates, data,
                        eps: float | None = None):
    """∇loglikelihood!(G, composite, models, data) -- src/fitting/fitting_base.jl:265-285.
    Fills G and overwrites ``composite`` with 1 - n/max(m, eps) (:274-280),
    eps = eps(eltype(composite)) (:277) unless given."""
    if not isinstance(models, HIPTemplates):
        raise TypeError("models must be a HIPTemplates")
    if G.size != models.ntemplates:
        raise ArgumentError("axes(G,1) == axes(models,2)")
    d = _as_data(data, models)
    if composite.size != models.nbins or d.nbins != models.nbins:
        raise ArgumentError("axes(models,1) == axes(data,1) == axes(composite,1)")
    e = _eps_of(_dtype_of(composite)) if eps is None else eps
    comp = _vec64(composite)
    g = np.empty(models.ntemplates, dtype=np.float64)
    models.ctx.check(models.ctx.lib.sfh_grad_loglikelihood_inplace(models.ctx.handle, models.handle,
                                                                   comp.ctypes.data, d.handle, e, g.ctypes.data))
    G[...] = g.reshape(G.shape)
    composite[...] = comp.reshape(composite.shape, order="F")
    return G


def fg_(F, G, coeffs, models: HIPTemplates, data, composite=None, eps: float | None = None,
        eps_grad: float | None = None):
    """fg!(F, G, coeffs, models, data, composite) -- src/fitting/solvers.jl:20-38.
    Returns -logL when F is not None; fills G (if not None) with -∇logL;
    ``composite`` (optional) is left as the reference leaves it.

    The reference's two clamps follow the composite's element type: the
    log-likelihood uses eps(promote_type(eltype(composite), eltype(data)))
    (fitting_base.jl:86) and ∇loglikelihood! eps(eltype(composite)) (:277).
    With a host ``composite`` both default to those; without one the composite
    is taken as FP64 (both eps(Float64)); the drivers pass
    ``driver_clamps(models, data)``."""
    if not isinstance(models, HIPTemplates):
        raise TypeError("models must be a HIPTemplates")
    N, T = models.shape
    c = _vec64(coeffs, T, "coeffs")
    d = _as_data(data, models)
    if d.nbins != N:
        raise ArgumentError("axes(models,1) == axes(data,1) == axes(composite,1)")
    if composite is not None and composite.size != N:
        raise ArgumentError("axes(models,1) == axes(data,1) == axes(composite,1)")
    if G is not None and G.size != T:
        raise ArgumentError("axes(G,1) == axes(models,2)")
    if composite is not None:
        e = _eps_of(composite, data) if eps is None else eps
        eg = _eps_of(_dtype_of(composite)) if eps_grad is None else eps_grad
    else:
        e = _f64eps if eps is None else eps
        eg = e if eps_grad is None else eps_grad
    want_F = F is not None
    if not want_F and G is None:
        if composite is not None:
            composite_(composite, c, models)
        return None
    negll = C.c_double()
    g = np.empty(T, dtype=np.float64) if G is not None else None
    comp = np.empty(N, dtype=np.float64) if composite is not None else None
    models.ctx.check(models.ctx.lib.sfh_fg(models.ctx.handle, models.handle, d.handle, c.ctypes.data,
                                           1 if want_F else 0, e, eg, C.byref(negll),
                                           g.ctypes.data if g is not None else None,
                                           comp.ctypes.data if comp is not None else None))
    if G is not None:
        G[...] = g.reshape(G.shape)
    if composite is not None:
        composite[...] = comp.reshape(composite.shape, order="F")
    if want_F:
        return negll.value
    return G


def fg_device(models: HIPTemplates, data: HIPData, d_coeffs, d_negLL=None, d_G=None, d_composite=None,
              want_F: bool = True, eps: float = _f64eps, eps_grad: float | None = None):
    """Asynchronous fg! on device buffers (torch tensors or raw pointers);
    ``eps_grad`` defaults to ``eps``."""
    eg = eps if eps_grad is None else eps_grad
    models.ctx.check(models.ctx.lib.sfh_fg_device(models.ctx.handle, models.handle, data.handle,
                                                  _ptr(d_coeffs), 1 if want_F else 0, eps, eg, _ptr(d_negLL),
                                                  _ptr(d_G), _ptr(d_composite)))


def fg_device_noalloc(models: HIPTemplates, data: HIPData, d_coeffs, d_negLL=None, d_G=None, d_composite=None,
                      want_F: bool = True, eps: float = _f64eps, eps_grad: float | None = None):
    """fg_device through sfh_fg_device_noalloc: the same kernels and results
    with the partial sums in the library's static device buffer, so the call
    allocates nothing (graph capture, serving loops).  At most 704 templates;
    calls that overlap in time on one GPU race."""
    eg = eps if eps_grad is None else eps_grad
    models.ctx.check(models.ctx.lib.sfh_fg_device_noalloc(models.ctx.handle, models.handle, data.handle,
                                                          _ptr(d_coeffs), 1 if want_F else 0, eps, eg,
                                                          _ptr(d_negLL), _ptr(d_G), _ptr(d_composite)))


def fg_batched(X, models: HIPTemplates, data, precision: str = "f64", want_G: bool = True,
               eps: float = _f64eps, eps_grad: float | None = None):
    """B independent fg! evaluations (columns of X, ntemplates x B) -- the
    HMC / ensemble call pattern (src/fitting/hmc_sample.jl:24-37,
    src/fitting/mcmc_sample.jl:12-23).  Returns (negLL[B], G[T,B] or None).
    ``eps_grad`` (the gradient's clamp) defaults to ``eps``."""
    eg = eps if eps_grad is None else eps_grad
    T = models.ntemplates
    Xf = np.asfortranarray(X, dtype=np.float64)
    if Xf.ndim != 2 or Xf.shape[0] != T:
        raise ArgumentError("X must be ntemplates x B")
    B = Xf.shape[1]
    d = _as_data(data, models)
    negll = np.empty(B, dtype=np.float64)
    G = np.empty((T, B), dtype=np.float64, order="F") if want_G else None
    models.ctx.check(models.ctx.lib.sfh_fg_batched(models.ctx.handle, models.handle, d.handle, Xf.ctypes.data,
                                                   B, _lib.PRECISIONS[precision], eps, eg, negll.ctypes.data,
                                                   G.ctypes.data if G is not None else None))
    return negll, G


def fg_batched_device(models: HIPTemplates, data: HIPData, d_X, B: int, d_negLL, d_G=None,
                      precision: str = "f64", eps: float = _f64eps, eps_grad: float | None = None):
    """Asynchronous batched fg! on device buffers (X: T x B column-major FP64)."""
    eg = eps if eps_grad is None else eps_grad
    models.ctx.check(models.ctx.lib.sfh_fg_batched_device(models.ctx.handle, models.handle, data.handle,
                                                          _ptr(d_X), int(B), _lib.PRECISIONS[precision], eps, eg,
                                                          _ptr(d_negLL), _ptr(d_G)))
