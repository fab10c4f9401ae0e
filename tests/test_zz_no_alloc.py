"""The allocation-free GEMV (zmlhip_dgemv_noalloc) against zmlhip_dgemv on the
GPU, through the C ABI: two calls in a row (the second smaller, so the shared
static scratch holds the first call's partial sums when it starts), every
layout/transpose form, bit-identical results; a size beyond the static array
is rejected before anything is queued. Runs alone and after the other tests.
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import COL, ROW, P, store

pytestmark = pytest.mark.gpu
ENOTSUP = -1003


def _call(zh, torch, fn, order, trans, m, n, alpha, A, x, incx, beta, y0, incy):
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lda = (m if order == COL else n) + 3
    a = torch.from_numpy(store(order, A, lda)).cuda()
    xd = torch.from_numpy(x.copy()).cuda()
    yd = torch.from_numpy(y0.copy()).cuda()
    rc = fn(order, trans, m, n, alpha, P(a), lda, P(xd), incx, beta, P(yd), incy, s)
    torch.cuda.synchronize()
    return rc, yd.cpu().numpy()


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("trans", [111, 112])
def test_noalloc_gemv_matches_dgemv_twice(zh, torch_cuda, order, trans):
    torch = torch_cuda
    rng = np.random.default_rng(order + trans)
    for m, n, incx, incy in ((5000, 3000, 1, -2), (1000, 700, -3, 1)):   # the second call is smaller
        A = rng.uniform(-1, 1, (m, n))
        lx, ly = (n, m) if trans == 111 else (m, n)
        x = rng.uniform(-1, 1, abs(incx) * (lx - 1) + 1)
        y0 = rng.uniform(-1, 1, abs(incy) * (ly - 1) + 1)
        rc1, y_ref = _call(zh, torch, zh.zmlhip_dgemv, order, trans, m, n, 0.75, A, x, incx, -1.5, y0, incy)
        rc2, y_new = _call(zh, torch, zh.zmlhip_dgemv_noalloc, order, trans, m, n, 0.75, A, x, incx, -1.5, y0, incy)
        assert rc1 == rc2 == 0
        assert np.array_equal(y_new, y_ref)   # same kernels, same sums, same order
        Aop = A if trans == 111 else A.T
        xv = x[::abs(incx)][::1 if incx > 0 else -1]
        yv = y0[::abs(incy)][::1 if incy > 0 else -1]
        want = 0.75 * (Aop @ xv) - 1.5 * yv
        got = y_new[::abs(incy)][::1 if incy > 0 else -1]
        assert np.abs(got - want).max() <= 1e-12 * (np.abs(Aop) @ np.abs(xv)).max()


def test_noalloc_gemv_rejects_more_rows_than_its_scratch(zh, torch_cuda):
    torch = torch_cuda
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    m = (1 << 20) + 1
    y = torch.zeros(1, dtype=torch.float64, device="cuda")
    # rejected on the host before any pointer is read: no device memory is touched
    rc = zh.zmlhip_dgemv_noalloc(COL, 111, m, 1, 1.0, P(y), m, P(y), 1, 0.0, P(y), 1, s)
    assert rc == ENOTSUP
    assert "scratch" in zh.zmlhip_last_error().decode()
    assert zh.zmlhip_dgemv_noalloc(COL, 111, 4, 4, 1.0, P(y), 3, P(y), 1, 0.0, P(y), 1, s) == -7
