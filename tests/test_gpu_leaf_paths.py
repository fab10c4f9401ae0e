"""The two multi-workgroup LU leaf kernels of libzmlhip.so, through the C ABI.

With zmlhip_set_leaf_lookahead(1) (the default) a leaf publishes column k + 1's
pivot candidates before it finishes column k's update, and every reader
completes the received rows with that update itself; with 0 it finishes each
column first. The arithmetic of every element is the same, so both give the
same factors, pivots and info bit for bit, for FP64 (dgetrf, both panel
strategies) and FP32 (sgetrf), on random matrices, with exact ties, with a
zero column and with a NaN. At n = 700 zml's pivots and info are checked
against the oracle as well.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, ROW, P

pytestmark = pytest.mark.gpu


@pytest.fixture
def leaf_paths(zh):
    def run(fn):
        prev = zh.zmlhip_set_leaf_lookahead(0)
        try:
            plain = fn()
            zh.zmlhip_set_leaf_lookahead(1)
            ahead = fn()
        finally:
            zh.zmlhip_set_leaf_lookahead(prev)
        return plain, ahead
    yield run
    zh.zmlhip_set_leaf_lookahead(1)


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _matrix(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(-1, 1, (n, n))
    if kind == "ties":   # few distinct magnitudes: many exact ties in every pivot search
        return rng.integers(-3, 4, (n, n)).astype(np.float64)
    if kind == "zero_column":
        a = rng.uniform(-1, 1, (n, n))
        a[:, n // 3] = 0.0
        return a
    a = rng.uniform(-1, 1, (n, n))   # "nan"
    a[n // 2, 5] = np.nan
    return a


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("n", [700, 2050])
@pytest.mark.parametrize("kind", ["uniform", "ties", "zero_column", "nan"])
def test_dgetrf_leaf_paths_bitwise(zh, torch_cuda, lu_panel, leaf_paths, order, n, kind):
    torch = torch_cuda
    a0 = _matrix(kind, n, n + order + len(kind))
    buf0 = (a0 if order == ROW else a0.T).reshape(-1).copy()
    s = _stream(torch)

    def run():
        a = torch.from_numpy(buf0.copy()).cuda()
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        assert zh.zmlhip_dgetrf(order, n, n, P(a), n, P(ipiv), P(info), s) == 0
        torch.cuda.synchronize()
        return a.cpu().numpy(), ipiv.cpu().numpy(), int(info.item())

    (ap, pp, ip), (aa, pa, ia) = leaf_paths(run)
    assert ip == ia
    assert np.array_equal(pp, pa)
    assert np.array_equal(ap.view(np.uint64), aa.view(np.uint64))
    if kind != "nan" and n <= 1000:   # the oracle's scalar LU: small orders only
        ref = buf0.copy()
        ip_ref = np.zeros(n, np.int32)
        assert O.getrf(order, n, n, ref, n, ip_ref) == ia
        assert np.array_equal(pa, ip_ref)


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("n", [1100, 3000])
def test_sgetrf_leaf_paths_bitwise(zh, torch_cuda, leaf_paths, order, n):
    torch = torch_cuda
    rng = np.random.default_rng(n + order)
    buf0 = rng.uniform(-1, 1, n * n).astype(np.float32)
    s = _stream(torch)

    def run():
        a = torch.from_numpy(buf0.copy()).cuda()
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        assert zh.zmlhip_sgetrf(order, n, n, P(a), n, P(ipiv), P(info), s) == 0
        torch.cuda.synchronize()
        return a.cpu().numpy(), ipiv.cpu().numpy(), int(info.item())

    (ap, pp, ip), (aa, pa, ia) = leaf_paths(run)
    assert ip == ia == 0
    assert np.array_equal(pp, pa)
    assert np.array_equal(ap.view(np.uint32), aa.view(np.uint32))
