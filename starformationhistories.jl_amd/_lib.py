"""ctypes binding of libsfh_hip.so (declared in include/sfh.h).

The shared library is the only compute path of this package: if it is
missing, importing the package raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libsfh_hip.so")

SFH_OK = 0
SFH_ERR_ARGUMENT = -1
SFH_ERR_DEVICE = -2
SFH_ERR_ALLOC = -3
SFH_ERR_DOMAIN = -4

SFH_PREC_FP32 = 0
SFH_PREC_BF16X2 = 1
SFH_PREC_F64 = 2

PRECISIONS = {"fp32": SFH_PREC_FP32, "bf16x2": SFH_PREC_BF16X2, "f64": SFH_PREC_F64}


class ArgumentError(ValueError):
    """Shape / value check failed (the reference's @argcheck ArgumentError)."""


class InexactError(ValueError):
    """Input outside the reference's domain (e.g. round(Int, NaN))."""


class DomainError(ValueError):
    """An argument outside the function's domain (the reference's DomainError)."""


class DeviceError(RuntimeError):
    """HIP runtime, launch or allocation failure."""


class sfh_range(C.Structure):
    _fields_ = [("first", C.c_double), ("step", C.c_double), ("length", C.c_int64)]


class sfh_table(C.Structure):
    _fields_ = [("first", C.c_double), ("step", C.c_double), ("length", C.c_int64), ("values", C.c_void_p)]


class sfh_fit_options(C.Structure):
    _fields_ = [("transform", C.c_int), ("method", C.c_int), ("memory", C.c_int), ("steps_per_graph", C.c_int),
                ("g_abstol", C.c_double), ("max_iterations", C.c_int64), ("eps", C.c_double),
                ("eps_grad", C.c_double)]


class sfh_fit_results(C.Structure):
    _fields_ = [("mu_map", C.c_void_p), ("mu_mle", C.c_void_p), ("invH_map", C.c_void_p),
                ("invH_mle", C.c_void_p), ("fmin", C.c_void_p), ("iterations", C.c_void_p),
                ("evaluations", C.c_void_p), ("status", C.c_void_p), ("steps", C.c_int64)]


SFH_FIT_LOG = 0
SFH_FIT_SQRT = 1
class sfh_nn_wcopy(C.Structure):
    _fields_ = [("src", C.c_void_p), ("rows", C.c_int64), ("cols", C.c_int64), ("dst", C.c_void_p),
                ("dstT", C.c_void_p)]


class sfh_imf(C.Structure):
    _fields_ = [("nseg", C.c_int32), ("breaks", C.c_void_p), ("slopes", C.c_void_p), ("coefs", C.c_void_p)]


class sfh_fn(C.Structure):
    _fields_ = [("kind", C.c_int32), ("p", C.c_double * 4), ("table", sfh_table), ("powerlaw", sfh_imf)]


SFH_FN_CONST = 0
SFH_FN_EXP_PHOTERR = 1
SFH_FN_MARTIN2016 = 2
SFH_FN_TABLE = 3
SFH_FN_POWERLAW = 4


class sfh_binary_model(C.Structure):
    _fields_ = [("kind", C.c_int32), ("fraction", C.c_double), ("qmin", C.c_double), ("qmax", C.c_double)]


MOCK_MAX_FILTERS = 8

SFH_OPT_BFGS = 0
SFH_OPT_LBFGS = 1
FIT_STATUS = {0: "running", 1: "converged", 2: "max_iterations", 3: "linesearch_failed"}

_vp = C.c_void_p
_i64 = C.c_int64
_dbl = C.c_double
_int = C.c_int
_p64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)

# (name, restype, argtypes) for every symbol of include/sfh.h
SIGNATURES = [
    ("sfh_version", _int, []),
    ("sfh_ctx_create", _int, [_int, _vp, C.POINTER(_vp)]),
    ("sfh_ctx_destroy", _int, [_vp]),
    ("sfh_last_error", C.c_char_p, [_vp]),
    ("sfh_ctx_synchronize", _int, [_vp]),
    ("sfh_ctx_set_timing", _int, [_vp, _int]),
    ("sfh_ctx_kernel_stats", _int, [_vp, C.c_char_p, _pd, _p64]),
    ("sfh_probe_valu_issue", _int, [_vp, C.c_int64, _pd]),
    ("sfh_ctx_reset_stats", _int, [_vp]),
    ("sfh_templates_create", _int, [_vp, _vp, _i64, _i64, C.POINTER(_vp)]),
    ("sfh_templates_create_device", _int, [_vp, _vp, _i64, _i64, C.POINTER(_vp)]),
    ("sfh_templates_destroy", _int, [_vp]),
    ("sfh_templates_shape", _int, [_vp, _p64, _p64]),
    ("sfh_templates_device_ptr", _vp, [_vp]),
    ("sfh_data_create", _int, [_vp, _vp, _i64, C.POINTER(_vp)]),
    ("sfh_data_create_device", _int, [_vp, _vp, _i64, C.POINTER(_vp)]),
    ("sfh_data_destroy", _int, [_vp]),
    ("sfh_composite", _int, [_vp, _vp, _vp, _vp]),
    ("sfh_composite_device", _int, [_vp, _vp, _vp, _vp]),
    ("sfh_loglikelihood", _int, [_vp, _vp, _vp, _i64, _dbl, _pd]),
    ("sfh_loglikelihood_coeffs", _int, [_vp, _vp, _vp, _vp, _dbl, _pd]),
    ("sfh_grad_loglikelihood", _int, [_vp, _vp, _vp, _vp, _dbl, _vp]),
    ("sfh_grad_loglikelihood_inplace", _int, [_vp, _vp, _vp, _vp, _dbl, _vp]),
    ("sfh_fg", _int, [_vp, _vp, _vp, _vp, _int, _dbl, _dbl, _pd, _vp, _vp]),
    ("sfh_fg_device", _int, [_vp, _vp, _vp, _vp, _int, _dbl, _dbl, _vp, _vp, _vp]),
    ("sfh_fg_device_noalloc", _int, [_vp, _vp, _vp, _vp, _int, _dbl, _dbl, _vp, _vp, _vp]),
    ("sfh_fg_batched", _int, [_vp, _vp, _vp, _vp, _i64, _int, _dbl, _dbl, _vp, _vp]),
    ("sfh_fg_batched_device", _int, [_vp, _vp, _vp, _vp, _i64, _int, _dbl, _dbl, _vp, _vp]),
    ("sfh_multifit_create", _int, [_vp, _i64, _vp, _vp, C.POINTER(_vp)]),
    ("sfh_multifit_destroy", _int, [_vp]),
    ("sfh_multifit_get_shape", _int, [_vp, _p64, C.POINTER(_int), C.POINTER(_int)]),
    ("sfh_multifit_set_shape", _int, [_vp, _i64, _int, _int]),
    ("sfh_multifit_autotune", _int, [_vp, _int, _p64, C.POINTER(_int), C.POINTER(_int)]),
    ("sfh_multifit_fg", _int, [_vp, _vp, _dbl, _dbl, _vp, _vp]),
    ("sfh_multifit_fg_device", _int, [_vp, _vp, _dbl, _dbl, _vp, _vp]),
    ("sfh_fit_options_default", _int, [C.POINTER(sfh_fit_options)]),
    ("sfh_multifit_fit", _int, [_vp, _vp, C.POINTER(sfh_fit_options), C.POINTER(sfh_fit_results)]),
    ("sfh_binary_pairs_hist", _int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _int, _dbl, C.POINTER(sfh_table), C.POINTER(sfh_table),
      C.POINTER(sfh_table), C.POINTER(sfh_table), sfh_range, sfh_range, _vp]),
    ("sfh_binary_pairs_hist_batch", _int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _int, _vp, C.POINTER(sfh_table), C.POINTER(sfh_table),
      C.POINTER(sfh_table), C.POINTER(sfh_table), sfh_range, sfh_range, _vp]),
    ("sfh_bin_cmd_smooth", _int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _int, sfh_range, sfh_range, _vp]),
    ("sfh_generate_stars", _int, [_vp, _i64, _vp, C.c_int32, _vp, C.c_int32, _dbl, C.c_int32, _dbl, _dbl, C.c_int32,
                                  C.POINTER(sfh_imf), C.POINTER(sfh_binary_model), C.c_uint64, C.c_uint64,
                                  C.POINTER(_vp)]),
    ("sfh_catalog_from_host", _int, [_vp, _i64, C.c_int32, _vp, C.POINTER(_vp)]),
    ("sfh_catalog_shape", _int, [_vp, _p64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _p64]),
    ("sfh_catalog_copy", _int, [_vp, _vp, _vp, _vp]),
    ("sfh_catalog_range", _int, [_vp, _vp, _vp]),
    ("sfh_catalog_destroy", _int, [_vp]),
    ("sfh_model_cmd", _int, [_vp, _vp, C.POINTER(sfh_table), C.POINTER(sfh_table), C.POINTER(sfh_table),
                             C.c_uint64, C.c_uint64, C.POINTER(_vp)]),
    ("sfh_bin_cmd", _int, [_vp, _i64, _vp, _vp, _vp, sfh_range, sfh_range, _vp]),
    ("sfh_bin_cmd_catalog", _int, [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, sfh_range, sfh_range, _vp]),
    ("sfh_nn_gemm_nt", _int, [_vp, _int, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, C.c_float, C.c_float]),
    ("sfh_nn_bias_act", _int, [_vp, _int, _vp, _i64, _vp, _i64, _i64, _int, _vp, _vp, _i64]),
    ("sfh_nn_linear", _int, [_vp, _int, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _vp, _int, _vp, _vp, _vp, _i64,
                             _vp, _i64]),
    ("sfh_nn_act_bwd", _int, [_vp, _vp, _vp, _i64, _int, _vp]),
    ("sfh_nn_act_bwd_chansum", _int, [_vp, _vp, _vp, _i64, _int, _vp, _vp]),
    ("sfh_nn_transpose", _int, [_vp, _int, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _int]),
    ("sfh_nn_dense_bwd_prep", _int, [_vp, _int, _vp, _vp, _i64, _i64, _int, _vp, _vp, _vp, _i64]),
    ("sfh_nn_weight_copies", _int, [_vp, _int, _i64, _vp]),
    ("sfh_nn_colsum", _int, [_vp, _vp, _i64, _i64, _i64, C.c_float, _vp]),
    ("sfh_nn_col2im", _int, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp]),
    ("sfh_nn_im2col", _int, [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _i64]),
    ("sfh_nn_im2col_t", _int, [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _i64]),
    ("sfh_nn_conv_t_fwd", _int, [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    ("sfh_nn_conv_t_dweight", _int, [_vp, _int, _vp, _vp, _int, _i64, _i64, _i64, _i64, _vp]),
    ("sfh_nn_conv_t_dinput", _int, [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64]),
    ("sfh_nn_conv1_fwd", _int, [_vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _int, _vp]),
    ("sfh_nn_conv1_bwd", _int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _int, _vp, _vp, _vp, _vp]),
    ("sfh_nn_mse", _int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, C.POINTER(C.c_float), _vp]),
    ("sfh_nn_template_loss", _int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, _vp, _vp]),
    ("sfh_ctx_capture_begin", _int, [_vp]),
    ("sfh_ctx_capture_end", _int, [_vp, C.POINTER(_vp)]),
    ("sfh_graph_launch", _int, [_vp, _i64]),
    ("sfh_graph_destroy", _int, [_vp]),
    ("sfh_nn_select_rows", _int, [_vp, _vp, _vp, _i64, _vp]),
    ("sfh_nn_counter_add", _int, [_vp, _vp, _i64]),
    ("sfh_nn_adam_sched", _int, [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, C.c_float, C.c_float, C.c_float,
                                 C.c_float]),
    ("sfh_nn_gather_rows", _int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64]),
    ("sfh_nn_log_target", _int, [_vp, _vp, _i64, _i64, _i64, _i64, C.c_float, C.c_float, C.c_float, _vp]),
    ("sfh_nn_log_template", _int, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, C.c_float, C.c_float, C.c_float, _vp,
                                   _i64]),
    ("sfh_nn_template_kl", _int, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, C.c_double, _vp]),
    ("sfh_nn_adam", _int, [_vp, _vp, _vp, _vp, _vp, _i64, C.c_float, C.c_float, C.c_float, C.c_float, _i64,
                           C.c_float]),
    ("sfh_partial_cmd_smooth_batch", _int,
     [_vp, _i64, _vp, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp, C.POINTER(sfh_fn), _vp,
      _vp, _vp, _dbl, sfh_range, sfh_range, _vp, _vp, _i64, _p64, _p64]),
    ("sfh_buffer_alloc", _int, [_vp, _i64, C.POINTER(_vp)]),
    ("sfh_buffer_free", _int, [_vp, _vp]),
    ("sfh_buffer_upload", _int, [_vp, _vp, _vp, _i64]),
    ("sfh_buffer_download", _int, [_vp, _vp, _vp, _i64]),
    ("sfh_bin_cmd_smooth_batch_device", _int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, sfh_range, sfh_range, _vp, _i64, _p64]),
]

_lib = None


def load():
    """Load libsfh_hip.so and declare every entry point; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
    # (SONAME libamdhip64.so.7) and loads it under the name libamdhip64.so.
    # Loading torch first makes this library's libamdhip64.so.7 dependency
    # resolve to that same runtime instead of a second copy from /opt/rocm.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libsfh_hip.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc == SFH_OK:
        return
    msg = ""
    if _lib is not None:
        raw = _lib.sfh_last_error(ctx)
        msg = raw.decode() if raw else ""
    if rc == SFH_ERR_ARGUMENT:
        raise ArgumentError(msg or "ArgumentError")
    if rc == SFH_ERR_DOMAIN:
        raise InexactError(msg or "InexactError")
    raise DeviceError(f"sfh error {rc}: {msg}")
