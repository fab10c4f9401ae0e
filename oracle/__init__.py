"""CPU oracle for zml's dense factor-and-solve path -- TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes + numpy) for ``liboracle.so``, the C restatement of
zml's BLAS/LAPACK routines in ``zml_oracle.c``. Only the test suite, the
smoke check and the CPU-baseline leg of ``bench.py`` use this module; the
product (``zml_amd``) never imports it.

All matrices are numpy float64 arrays handed over as flat buffers together
with an explicit layout (101 row-major / 102 column-major) and leading
dimension, exactly like the C entry points.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ROW, COL = 101, 102
NOTRANS, TRANS, CONJTRANS, CONJNOTRANS = 111, 112, 113, 114
UPPER, LOWER = 121, 122
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142


def build(force: bool = True) -> str:
    """Compile liboracle.so with the committed Makefile."""
    cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        d, i, p, u64 = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64
        f = ctypes.c_float
        sig = {
            "zr_lamch": (d, [ctypes.c_char]),
            "zr_iamax": (i, [i, p, i]),
            "zr_scal": (i, [i, d, p, i]),
            "zr_swap": (i, [i, p, i, p, i]),
            "zr_gemm": (i, [i, i, i, i, i, i, d, p, i, p, i, d, p, i]),
            "zr_trsm": (i, [i, i, i, i, i, i, i, d, p, i, p, i]),
            "zr_syrk": (i, [i, i, i, i, i, d, p, i, d, p, i]),
            "zr_gemv": (i, [i, i, i, i, d, p, i, p, i, d, p, i]),
            "zr_trsv": (i, [i, i, i, i, i, p, i, p, i]),
            "zr_lange": (d, [i, ctypes.c_char, i, i, p, i]),
            "zr_slamch": (ctypes.c_float, [ctypes.c_char]),
            "zr_sgemm": (i, [i, i, i, i, i, i, f, p, i, p, i, f, p, i]),
            "zr_strsm": (i, [i, i, i, i, i, i, i, f, p, i, p, i]),
            "zr_ssyrk": (i, [i, i, i, i, i, f, p, i, f, p, i]),
            "zr_strsv": (i, [i, i, i, i, i, p, i, p, i]),
            "zr_sgetrf": (i, [i, i, i, p, i, p]),
            "zr_sgetrs": (i, [i, i, i, i, p, i, p, p, i]),
            "zr_spotrf": (i, [i, i, i, p, i]),
            "zr_spotrs": (i, [i, i, i, i, p, i, p, i]),
            "zr_sgetrf_unblocked": (i, [i, i, i, p, i, p]),
            "zr_spotrf_lower_unblocked": (i, [i, p, i]),
            "zr_set_trsm_faithful_lda": (None, [i]),
            "zr_laswp": (i, [i, i, p, i, i, i, p, i]),
            "zr_getrf2": (i, [i, i, i, p, i, p]),
            "zr_getrf": (i, [i, i, i, p, i, p]),
            "zr_getrs": (i, [i, i, i, i, p, i, p, p, i]),
            "zr_gesv": (i, [i, i, i, p, i, p, p, i]),
            "zr_potrf2": (i, [i, i, i, p, i]),
            "zr_potrf": (i, [i, i, i, p, i]),
            "zr_potrs": (i, [i, i, i, i, p, i, p, i]),
            "zr_posv": (i, [i, i, i, i, p, i, p, i]),
            "zr_getrf_unblocked": (i, [i, i, i, p, i, p]),
            "zr_getrf_fast": (i, [i, i, p, i, p, i]),
            "zr_potrf_lower_unblocked": (i, [i, p, i]),
            "zr_rng_u64": (u64, [u64, u64]),
            "zr_rng_uniform": (d, [u64, u64]),
            "zr_fill_uniform": (None, [i, i, i, p, i, u64, d, d]),
            "zr_norm_inf": (d, [i, i, i, p, i]),
            "zr_residual_inf": (d, [i, i, p, i, p, p, p]),
            "zr_backward_error": (d, [i, i, p, i, p, p]),
            "zr_min_pivot_gap": (d, [i, i, p, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    """Pointer to a C-contiguous numpy buffer (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    a = np.asarray(a)
    if a.dtype != np.float64:
        raise TypeError("oracle takes float64 buffers")
    return a


# ---- BLAS ------------------------------------------------------------------
def iamax(n, x, incx):
    return lib().zr_iamax(n, _p(_f64(x)), incx)


def scal(n, alpha, x, incx):
    return lib().zr_scal(n, alpha, _p(_f64(x)), incx)


def swap(n, x, incx, y, incy):
    return lib().zr_swap(n, _p(_f64(x)), incx, _p(_f64(y)), incy)


def gemm(order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    return lib().zr_gemm(order, ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def trsm(order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    return lib().zr_trsm(order, side, uplo, ta, diag, m, n, alpha, _p(a), lda, _p(b), ldb)


def syrk(order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    return lib().zr_syrk(order, uplo, trans, n, k, alpha, _p(a), lda, beta, _p(c), ldc)


def gemv(order, trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    return lib().zr_gemv(order, trans, m, n, alpha, _p(a), lda, _p(x), incx, beta, _p(y), incy)


def trsv(order, uplo, trans, diag, n, a, lda, x, incx):
    return lib().zr_trsv(order, uplo, trans, diag, n, _p(a), lda, _p(x), incx)


def lamch(c):
    return lib().zr_lamch(c.encode() if isinstance(c, str) else c)


def lange(order, norm, m, n, a, ld):
    return lib().zr_lange(order, norm.encode() if isinstance(norm, str) else norm, m, n, _p(a), ld)


# ---- f32 (zml's generic routines at T = f32) ----------------------------------
def _f32(a):
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise TypeError("the f32 oracle takes float32 buffers")
    return a


def slamch(c):
    return lib().zr_slamch(c.encode() if isinstance(c, str) else c)


def sgemm(order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    return lib().zr_sgemm(order, ta, tb, m, n, k, alpha, _p(_f32(a)), lda, _p(_f32(b)), ldb, beta,
                          _p(_f32(c)), ldc)


def strsm(order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    return lib().zr_strsm(order, side, uplo, ta, diag, m, n, alpha, _p(_f32(a)), lda, _p(_f32(b)), ldb)


def ssyrk(order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    return lib().zr_ssyrk(order, uplo, trans, n, k, alpha, _p(_f32(a)), lda, beta, _p(_f32(c)), ldc)


def strsv(order, uplo, trans, diag, n, a, lda, x, incx):
    return lib().zr_strsv(order, uplo, trans, diag, n, _p(_f32(a)), lda, _p(_f32(x)), incx)


def sgetrf(order, m, n, a, lda, ipiv):
    return lib().zr_sgetrf(order, m, n, _p(_f32(a)), lda, _p(ipiv))


def sgetrs(order, trans, n, nrhs, a, lda, ipiv, b, ldb):
    return lib().zr_sgetrs(order, trans, n, nrhs, _p(_f32(a)), lda, _p(ipiv), _p(_f32(b)), ldb)


def spotrf(order, uplo, n, a, lda):
    return lib().zr_spotrf(order, uplo, n, _p(_f32(a)), lda)


def spotrs(order, uplo, n, nrhs, a, lda, b, ldb):
    return lib().zr_spotrs(order, uplo, n, nrhs, _p(_f32(a)), lda, _p(_f32(b)), ldb)


def sgetrf_unblocked(order, m, n, a, lda, ipiv):
    return lib().zr_sgetrf_unblocked(order, m, n, _p(_f32(a)), lda, _p(ipiv))


def spotrf_lower_unblocked(n, a, lda):
    return lib().zr_spotrf_lower_unblocked(n, _p(_f32(a)), lda)


def set_trsm_faithful_lda(on: bool):
    lib().zr_set_trsm_faithful_lda(1 if on else 0)


# ---- LAPACK ----------------------------------------------------------------
def laswp(order, n, a, lda, k1, k2, ipiv, incx):
    return lib().zr_laswp(order, n, _p(a), lda, k1, k2, _p(ipiv), incx)


def getrf2(order, m, n, a, lda, ipiv):
    return lib().zr_getrf2(order, m, n, _p(a), lda, _p(ipiv))


def getrf(order, m, n, a, lda, ipiv):
    return lib().zr_getrf(order, m, n, _p(a), lda, _p(ipiv))


def getrs(order, trans, n, nrhs, a, lda, ipiv, b, ldb):
    return lib().zr_getrs(order, trans, n, nrhs, _p(a), lda, _p(ipiv), _p(b), ldb)


def gesv(order, n, nrhs, a, lda, ipiv, b, ldb):
    return lib().zr_gesv(order, n, nrhs, _p(a), lda, _p(ipiv), _p(b), ldb)


def potrf2(order, uplo, n, a, lda):
    return lib().zr_potrf2(order, uplo, n, _p(a), lda)


def potrf(order, uplo, n, a, lda):
    return lib().zr_potrf(order, uplo, n, _p(a), lda)


def potrs(order, uplo, n, nrhs, a, lda, b, ldb):
    return lib().zr_potrs(order, uplo, n, nrhs, _p(a), lda, _p(b), ldb)


def posv(order, uplo, n, nrhs, a, lda, b, ldb):
    return lib().zr_posv(order, uplo, n, nrhs, _p(a), lda, _p(b), ldb)


def getrf_unblocked(order, m, n, a, lda, ipiv):
    return lib().zr_getrf_unblocked(order, m, n, _p(a), lda, _p(ipiv))


def getrf_fast(m, n, a, lda, ipiv, nthreads=0):
    return lib().zr_getrf_fast(m, n, _p(a), lda, _p(ipiv), nthreads)


def potrf_lower_unblocked(n, a, lda):
    return lib().zr_potrf_lower_unblocked(n, _p(a), lda)


# ---- synthetic inputs and metrics -------------------------------------------
def rng_u64(seed, ctr):
    return lib().zr_rng_u64(seed, ctr)


def fill_uniform(order, m, n, seed, lo=-1.0, hi=1.0, ld=None):
    """Return a flat float64 buffer holding the seeded m x n matrix."""
    ld = ld if ld is not None else (m if order == COL else n)
    rows_or_cols = n if order == COL else m
    a = np.zeros(ld * rows_or_cols, dtype=np.float64)
    lib().zr_fill_uniform(order, m, n, _p(a), ld, seed, lo, hi)
    return a


def norm_inf(order, m, n, a, ld):
    return lib().zr_norm_inf(order, m, n, _p(a), ld)


def backward_error(order, n, a, lda, x, b):
    return lib().zr_backward_error(order, n, _p(a), lda, _p(x), _p(b))


def min_pivot_gap(order, n, a, lda):
    return lib().zr_min_pivot_gap(order, n, _p(a), lda)
