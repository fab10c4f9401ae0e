"""ctypes binding of libzmlhip.so (the C ABI declared in include/*.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzmlhip.so")

_c_int, _c_double, _p, _c_char = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char
_c_float, _c_long = ctypes.c_float, ctypes.c_long

# name -> (restype, argtypes); mirrors include/cblas.h, clapack.h, zmlhip.h
SIGNATURES = {
    # cblas.h
    "cblas_dscal": (None, [_c_int, _c_double, _p, _c_int]),
    "cblas_dswap": (None, [_c_int, _p, _c_int, _p, _c_int]),
    "cblas_idamax": (ctypes.c_uint, [_c_int, _p, _c_int]),
    "cblas_dgemv": (None, [_c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int, _p, _c_int,
                           _c_double, _p, _c_int]),
    "cblas_dgemm": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int,
                           _p, _c_int, _c_double, _p, _c_int]),
    "cblas_dsyrk": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int,
                           _c_double, _p, _c_int]),
    "cblas_dtrsm": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p,
                           _c_int, _p, _c_int]),
    "cblas_dtrsv": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _p, _c_int, _p, _c_int]),
    "cblas_sgemm": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p, _c_int,
                           _p, _c_int, _c_float, _p, _c_int]),
    "cblas_ssyrk": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p, _c_int,
                           _c_float, _p, _c_int]),
    "cblas_strsm": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p,
                           _c_int, _p, _c_int]),
    # clapack.h
    "LAPACKE_dlamch": (_c_double, [_c_char]),
    "LAPACKE_dlaswp": (_c_int, [_c_int, _c_int, _p, _c_int, _c_int, _c_int, _p, _c_int]),
    "LAPACKE_dgetrf2": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p]),
    "LAPACKE_dgetrf": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p]),
    "LAPACKE_dgetrs": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _p, _c_int]),
    "LAPACKE_dgesv": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p, _p, _c_int]),
    "LAPACKE_dpotrf2": (_c_int, [_c_int, _c_char, _c_int, _p, _c_int]),
    "LAPACKE_dpotrf": (_c_int, [_c_int, _c_char, _c_int, _p, _c_int]),
    "LAPACKE_dpotrs": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _c_int]),
    "LAPACKE_dposv": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _c_int]),
    "LAPACKE_sgetrf": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p]),
    "LAPACKE_spotrf": (_c_int, [_c_int, _c_char, _c_int, _p, _c_int]),
    "LAPACKE_dlange": (_c_double, [_c_int, _c_char, _c_int, _c_int, _p, _c_int]),
    # zmlhip.h
    "zmlhip_version": (_c_int, []),
    "zmlhip_last_error": (ctypes.c_char_p, []),
    "zmlhip_stream_synchronize": (_c_int, [_p]),
    "zmlhip_profile_enable": (_c_int, [_c_int]),
    "zmlhip_profile_collect": (_c_int, [_c_int, _p, _p, _p, _p]),
    "zmlhip_dgemm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int,
                              _p, _c_int, _c_double, _p, _c_int, _p]),
    "zmlhip_dtrsm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p,
                              _c_int, _p, _c_int, _p]),
    "zmlhip_dsyrk": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int,
                              _c_double, _p, _c_int, _p]),
    "zmlhip_idamax": (_c_int, [_c_int, _p, _c_int, _p, _p]),
    "zmlhip_dscal": (_c_int, [_c_int, _c_double, _p, _c_int, _p]),
    "zmlhip_dswap": (_c_int, [_c_int, _p, _c_int, _p, _c_int, _p]),
    "zmlhip_dgemv": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int, _p, _c_int,
                              _c_double, _p, _c_int, _p]),
    "zmlhip_dgemv_noalloc": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_double, _p, _c_int, _p, _c_int,
                                      _c_double, _p, _c_int, _p]),
    "zmlhip_dtrsv": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _p, _c_int, _p, _c_int, _p]),
    "zmlhip_sgemm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p, _c_int,
                              _p, _c_int, _c_float, _p, _c_int, _p]),
    "zmlhip_strsm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p,
                              _c_int, _p, _c_int, _p]),
    "zmlhip_ssyrk": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _p, _c_int,
                              _c_float, _p, _c_int, _p]),
    "zmlhip_dlange": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _p]),
    "zmlhip_sgetrf": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p, _p, _p]),
    "zmlhip_spotrf": (_c_int, [_c_int, _c_char, _c_int, _p, _c_int, _p, _p]),
    "zmlhip_set_panel_spin_limit": (_c_int, [_c_long]),
    "zmlhip_set_panel_one_launch_rows": (_c_int, [_c_long]),
    "zmlhip_set_gemm_lds_dma": (_c_int, [_c_int]),
    "zmlhip_set_leaf_lookahead": (_c_int, [_c_int]),
    "zmlhip_dlaswp": (_c_int, [_c_int, _c_int, _p, _c_int, _c_int, _c_int, _p, _c_int, _p]),
    "zmlhip_dgetrf": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p, _p, _p]),
    "zmlhip_dgetrs": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _p, _c_int, _p]),
    "zmlhip_dpotrf": (_c_int, [_c_int, _c_char, _c_int, _p, _c_int, _p, _p]),
    "zmlhip_dpotrs": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _c_int, _p]),
    "zmlhip_dgesv_mixed": (_c_int, [_c_int, _c_int, _c_int, _p, _c_int, _p, _c_int, _p, _c_int, _p, _p,
                                    _p]),
    "zmlhip_dposv_mixed": (_c_int, [_c_int, _c_char, _c_int, _c_int, _p, _c_int, _p, _c_int, _p, _c_int,
                                    _p, _p, _p]),
    "zmlhip_batch_plan_create": (_c_int, [_p, _c_int, _p, _p, _p, _p, _p, _p, _p, _c_int, _p]),
    "zmlhip_batch_plan_execute": (_c_int, [_p, _p]),
    "zmlhip_batch_plan_status": (_c_int, [_p, _p, _p]),
    "zmlhip_batch_plan_shapes": (_c_int, [_p, _p, _p, _p]),
    "zmlhip_batch_plan_destroy": (_c_int, [_p]),
}


class IROpts(ctypes.Structure):
    """zmlhip_ir_opts (include/zmlhip.h)."""
    _fields_ = [("precision", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("gmres_restart", ctypes.c_int), ("gmres_max_iter", ctypes.c_int),
                ("no_gmres", ctypes.c_int), ("tol", ctypes.c_double), ("no_optimistic", ctypes.c_int),
                ("pivot_threshold", ctypes.c_double)]


class IRResult(ctypes.Structure):
    """zmlhip_ir_result (include/zmlhip.h)."""
    _fields_ = [("info", ctypes.c_int), ("iterations", ctypes.c_int),
                ("gmres_iterations", ctypes.c_int), ("fallback", ctypes.c_int),
                ("backward_error", ctypes.c_double), ("panels", ctypes.c_int),
                ("panels_optimistic", ctypes.c_int)]

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libzmlhip.so; raises if it is missing or lacks a declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m zml_amd.build` (hipcc, gfx950)")
    # One HIP runtime per process: when PyTorch is present it brings its own
    # libamdhip64 (same soname); loading it first makes this library bind to
    # that copy instead of starting a second runtime on the same device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)  # AttributeError when a symbol is not exported
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().zmlhip_last_error().decode()
