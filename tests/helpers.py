"""Shared helpers of the test suite (pointer passing, layouts, tolerances)."""
import ctypes

import numpy as np

ROW, COL = 101, 102
U64 = 2.0 ** -53


def P(a):
    """ctypes pointer to a numpy buffer or a torch tensor (device pointer)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def dense(order, m, n, buf, ld):
    """Logical m x n view (copy) of a flat buffer stored in `order` with ld."""
    buf = np.asarray(buf)
    if m == 0 or n == 0:
        return np.zeros((m, n))
    if order == COL:
        need = ld * (n - 1) + m
        full = np.zeros(ld * n)
        full[:need] = buf[:need]
        return full.reshape(n, ld)[:, :m].T.copy()
    need = ld * (m - 1) + n
    full = np.zeros(ld * m)
    full[:need] = buf[:need]
    return full.reshape(m, ld)[:, :n].copy()


def store(order, M, ld, fill=0.0, rows_alloc=None):
    """Flat buffer holding the 2-D array M in `order` with leading dimension ld."""
    m, n = M.shape
    if order == COL:
        buf = np.full(ld * max(n, 1), fill)
        v = buf.reshape(max(n, 1), ld)
        v[:n, :m] = M.T
    else:
        buf = np.full(ld * max(m, 1), fill)
        v = buf.reshape(max(m, 1), ld)
        v[:m, :n] = M
    return buf


def rng_matrix(m, n, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, size=(m, n))
