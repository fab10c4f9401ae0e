"""The two FP64 GEMM kernels of libzmlhip.so, through the C ABI.

FP64 products with whole 16-wide k-slabs, 16-byte aligned operand columns and
an even extent along each operand's contiguous m / n direction move their
operand slabs global -> LDS with the load unit (gemm_dma_f64_kernel); the rest,
and every product while zmlhip_set_gemm_lds_dma(0) is in force, stage them
through registers (gemm_kernel). Both issue the same MFMAs in the same order,
so:

* every transpose / layout / alpha-beta form gives the same bits on both paths,
  and lies within the elementwise bound
  |C - C_ref| <= 2 (k + 2) u (|alpha| |A||B| + |beta| |C|)
  of numpy's float64 product (an independent library);
* triangular updates (dsyrk) and the FP64 factorizations built on these GEMMs
  (dgetrf, dpotrf) give the same bits on both paths.
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import COL, ROW, U64, P, dense, store

pytestmark = pytest.mark.gpu


@pytest.fixture
def paths(zh):
    """Runs fn() on the register-staged path, then on the LDS-DMA path (the
    default, restored afterwards); returns both results."""
    def run(fn):
        prev = zh.zmlhip_set_gemm_lds_dma(0)
        try:
            reg = fn()
            zh.zmlhip_set_gemm_lds_dma(1)
            dma = fn()
        finally:
            zh.zmlhip_set_gemm_lds_dma(prev)
        return reg, dma
    yield run
    zh.zmlhip_set_gemm_lds_dma(1)


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# (m, n, k, pad): the first four take the LDS-DMA path in every transpose form
# and layout (even extents and leading dimensions, k a multiple of 16); the last
# two fall back to the register path (odd extents and leading dimensions; k not
# a multiple of 16)
SHAPES = [(256, 256, 64, 0), (130, 98, 48, 2), (514, 300, 256, 2), (1000, 778, 512, 4),
          (131, 97, 48, 1), (256, 256, 40, 0)]
FORMS = [(-1.0, 1.0), (1.0, 1.0), (0.5, -2.0), (2.0, 0.0)]


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("ta", [111, 112])
@pytest.mark.parametrize("tb", [111, 112])
@pytest.mark.parametrize("m,n,k,pad", SHAPES)
@pytest.mark.parametrize("alpha,beta", FORMS)
def test_dgemm_paths_bitwise_and_bounded(zh, torch_cuda, paths, order, ta, tb, m, n, k, pad, alpha, beta):
    torch = torch_cuda
    rng = np.random.default_rng(m * 31 + n * 7 + k + ta + 2 * tb + order)
    Am = rng.uniform(-1, 1, (m, k)) if ta == 111 else rng.uniform(-1, 1, (k, m))
    Bm = rng.uniform(-1, 1, (k, n)) if tb == 111 else rng.uniform(-1, 1, (n, k))
    Cm = rng.uniform(-1, 1, (m, n))
    lda = (Am.shape[0] if order == COL else Am.shape[1]) + pad
    ldb = (Bm.shape[0] if order == COL else Bm.shape[1]) + pad
    ldc = (m if order == COL else n) + pad
    a = torch.from_numpy(store(order, Am, lda)).cuda()
    b = torch.from_numpy(store(order, Bm, ldb)).cuda()
    c0 = store(order, Cm, ldc, fill=np.nan)
    s = _stream(torch)

    def run():
        c = torch.from_numpy(c0.copy()).cuda()
        assert zh.zmlhip_dgemm(order, ta, tb, m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc, s) == 0
        torch.cuda.synchronize()
        return c.cpu().numpy()

    reg, dma = paths(run)
    assert np.array_equal(reg.view(np.uint64), dma.view(np.uint64))
    got = dense(order, m, n, dma, ldc)
    Aop = Am if ta == 111 else Am.T
    Bop = Bm if tb == 111 else Bm.T
    ref = alpha * (Aop @ Bop) + beta * Cm
    bound = 2 * (k + 2) * U64 * (abs(alpha) * (np.abs(Aop) @ np.abs(Bop)) + abs(beta) * np.abs(Cm)) + 1e-300
    assert np.all(np.abs(got - ref) <= 2 * bound)
    assert np.array_equal(np.isnan(dma), np.isnan(c0))   # padding never written


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("n,k,alpha,beta", [(300, 64, -1.0, 1.0), (1030, 128, 0.7, 0.0)])
def test_dsyrk_paths_bitwise(zh, torch_cuda, paths, order, uplo, trans, n, k, alpha, beta):
    torch = torch_cuda
    rng = np.random.default_rng(n + k + uplo + trans + order)
    Am = rng.uniform(-1, 1, (n, k)) if trans == 111 else rng.uniform(-1, 1, (k, n))
    Cm = rng.uniform(-1, 1, (n, n))
    lda = Am.shape[0] if order == COL else Am.shape[1]
    a = torch.from_numpy(store(order, Am, lda)).cuda()
    c0 = store(order, Cm, n)
    s = _stream(torch)

    def run():
        c = torch.from_numpy(c0.copy()).cuda()
        assert zh.zmlhip_dsyrk(order, uplo, trans, n, k, alpha, P(a), lda, beta, P(c), n, s) == 0
        torch.cuda.synchronize()
        return c.cpu().numpy()

    reg, dma = paths(run)
    assert np.array_equal(reg.view(np.uint64), dma.view(np.uint64))
    got = dense(order, n, n, dma, n)
    Aop = Am if trans == 111 else Am.T
    tri = np.tril if uplo == 122 else np.triu
    ref = alpha * (Aop @ Aop.T) + beta * Cm
    bound = 2 * (k + 2) * U64 * (abs(alpha) * (np.abs(Aop) @ np.abs(Aop).T) + abs(beta) * np.abs(Cm)) + 1e-300
    assert np.all(tri(np.abs(got - ref) <= 2 * bound) | (tri(np.ones((n, n))) == 0))
    other = np.triu(np.ones((n, n)), 1) if uplo == 122 else np.tril(np.ones((n, n)), -1)
    assert np.array_equal(got[other == 1], Cm[other == 1])


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("n", [1024, 2050])
def test_dgetrf_paths_bitwise(zh, torch_cuda, paths, order, n):
    torch = torch_cuda
    rng = np.random.default_rng(n + order)
    a0 = rng.uniform(-1, 1, n * n)
    s = _stream(torch)

    def run():
        a = torch.from_numpy(a0.copy()).cuda()
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        assert zh.zmlhip_dgetrf(order, n, n, P(a), n, P(ipiv), P(info), s) == 0
        torch.cuda.synchronize()
        return a.cpu().numpy(), ipiv.cpu().numpy(), int(info.item())

    (ar, pr, ir), (ad, pd, idd) = paths(run)
    assert ir == idd == 0
    assert np.array_equal(pr, pd)
    assert np.array_equal(ar.view(np.uint64), ad.view(np.uint64))


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dpotrf_paths_bitwise(zh, torch_cuda, paths, order, uplo):
    torch = torch_cuda
    n = 1536
    rng = np.random.default_rng(n + order + ord(uplo))
    G = rng.uniform(-1, 1, (n, n))
    a0 = (G @ G.T / n + np.eye(n)).reshape(-1)
    s = _stream(torch)

    def run():
        a = torch.from_numpy(a0.copy()).cuda()
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        assert zh.zmlhip_dpotrf(order, uplo.encode(), n, P(a), n, P(info), s) == 0
        torch.cuda.synchronize()
        return a.cpu().numpy(), int(info.item())

    (ar, ir), (ad, idd) = paths(run)
    assert ir == idd == 0
    assert np.array_equal(ar.view(np.uint64), ad.view(np.uint64))
