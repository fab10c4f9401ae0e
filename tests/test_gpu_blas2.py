"""GPU parity of the BLAS/LAPACK entries added beside the FP64 Level-3 set,
through the C ABI (libzmlhip.so):

* cblas_dgemv and cblas_dtrsv: the reference's own vectors
  (test/linalg/blas/gemv.zig, trsv.zig) replayed exactly as zml's tests state
  them (exact for gemv, rel 1e-7 for trsv), then every uplo/trans/diag/layout
  and stride sign against the oracle's restatement of trsv.zig;
* the FP32 entries cblas_sgemm/ssyrk/strsm: the reference's integer-valued
  gemm/syrk vectors exactly, the trsm vectors at f32 precision, and random
  operands against the oracle's f32 instantiation with an f32 gamma_k bound;
* LAPACKE_dlange / zmlhip_dlange against numpy's norms (an independent
  library) and the oracle;
* the FP64 square root of the Cholesky diagonal kernel is IEEE's: a diagonal
  SPD matrix factors to exactly numpy's sqrt of its diagonal.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests import fixture_replay
from tests.helpers import COL, ROW, U64, P, dense, store

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24


class CblasL2:
    def __init__(self, zh):
        self.zh = zh

    def gemv(self, order, trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
        self.zh.cblas_dgemv(order, trans, m, n, alpha, P(a), lda, P(x), incx, beta, P(y), incy)

    def trsv(self, order, uplo, trans, diag, n, a, lda, x, incx):
        self.zh.cblas_dtrsv(order, uplo, trans, diag, n, P(a), lda, P(x), incx)


@pytest.mark.parametrize("name", ["gemv", "trsv"])
def test_reference_level2_vectors_through_cblas(zh, name):
    fails = fixture_replay.replay(name, CblasL2(zh))
    assert not fails, "\n".join(fails[:10])


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112, 113])
@pytest.mark.parametrize("diag", [131, 132])
@pytest.mark.parametrize("n,incx", [(1, 1), (45, 1), (300, -2), (777, 3)])
def test_dtrsv_matches_oracle(zh, order, uplo, trans, diag, n, incx):
    rng = np.random.default_rng(n * 7 + incx + uplo + trans + diag + order)
    M = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    lda = n + 2
    a = store(order, M, lda)
    x = rng.uniform(-1, 1, abs(incx) * (n - 1) + 1)
    xr = x.copy()
    assert O.trsv(order, uplo, trans, diag, n, a, lda, xr, incx) == 0
    zh.cblas_dtrsv(order, uplo, trans, diag, n, P(a), lda, P(x), incx)
    # backward-stable substitution: |x - x_ref| within a small multiple of n u |x|
    assert np.all(np.abs(x - xr) <= 64 * n * U64 * (np.abs(xr).max() + 1.0)), np.abs(x - xr).max()


def test_dtrsv_device_api_and_invalid_arguments(zh, torch_cuda):
    torch = torch_cuda
    n = 64
    rng = np.random.default_rng(3)
    T = np.tril(rng.uniform(-1, 1, (n, n))) + 8 * np.eye(n)
    b = rng.uniform(-1, 1, n)
    a = torch.from_numpy(store(COL, T, n)).cuda()
    x = torch.from_numpy(b.copy()).cuda()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert zh.zmlhip_dtrsv(COL, 122, 111, 131, n, P(a), n, P(x), 1, s) == 0
    torch.cuda.synchronize()
    assert np.allclose(x.cpu().numpy(), np.linalg.solve(T, b), rtol=1e-12, atol=1e-14)
    assert zh.zmlhip_dtrsv(7, 122, 111, 131, n, P(a), n, P(x), 1, s) == -1
    assert zh.zmlhip_dtrsv(COL, 122, 111, 131, n, P(a), n - 1, P(x), 1, s) == -7
    assert zh.zmlhip_dtrsv(COL, 122, 111, 131, n, P(a), n, P(x), 0, s) == -9


# ---- FP32 entries -------------------------------------------------------------
class CblasS:
    def __init__(self, zh):
        self.zh = zh

    def gemm(self, order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        self.zh.cblas_sgemm(order, ta, tb, m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc)

    def trsm(self, order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
        self.zh.cblas_strsm(order, side, uplo, ta, diag, m, n, alpha, P(a), lda, P(b), ldb)

    def syrk(self, order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
        self.zh.cblas_ssyrk(order, uplo, trans, n, k, alpha, P(a), lda, beta, P(c), ldc)


@pytest.mark.parametrize("name", ["gemm", "syrk"])
def test_f32_reference_vectors_exact(zh, name):
    fails = fixture_replay.replay(name, CblasS(zh), dtype=np.float32)
    assert not fails, "\n".join(fails[:10])


def test_f32_trsm_reference_vectors(zh):
    """The chained f64 vectors at f32 precision. The blocked GPU solve (MFMA
    updates with fused multiply-add) rounds differently from the oracle's
    loops; the longest chain (trsm.zig:841-864) amplifies the f32 rounding to
    1.6e-4 relative there (4e-5 for the oracle), so the floor is 5e-4."""
    fails = fixture_replay.replay("trsm", CblasS(zh), dtype=np.float32, rel_floor=5e-4)
    assert not fails, "\n".join(fails[:10])


def _f32(M):
    return M.astype(np.float32)


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("ta", [111, 112])
@pytest.mark.parametrize("tb", [111, 112])
@pytest.mark.parametrize("m,n,k,alpha,beta", [(37, 53, 29, -1.0, 1.0), (130, 257, 70, 0.5, -2.0),
                                              (256, 128, 300, 1.0, 0.0)])
def test_sgemm_matches_f32_oracle(zh, order, ta, tb, m, n, k, alpha, beta):
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    Am = rng.uniform(-1, 1, (m, k)) if ta == 111 else rng.uniform(-1, 1, (k, m))
    Bm = rng.uniform(-1, 1, (k, n)) if tb == 111 else rng.uniform(-1, 1, (n, k))
    Cm = rng.uniform(-1, 1, (m, n))
    lda = (Am.shape[0] if order == COL else Am.shape[1]) + 1
    ldb = (Bm.shape[0] if order == COL else Bm.shape[1]) + 1
    ldc = (m if order == COL else n) + 1
    a, b = _f32(store(order, Am, lda)), _f32(store(order, Bm, ldb))
    c = _f32(store(order, Cm, ldc))
    c_ref = c.copy()
    assert O.sgemm(order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c_ref, ldc) == 0
    zh.cblas_sgemm(order, ta, tb, m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc)
    got = dense(order, m, n, c.astype(np.float64), ldc)
    ref = dense(order, m, n, c_ref.astype(np.float64), ldc)
    Aop = dense(order, *Am.shape, a.astype(np.float64), lda)
    Bop = dense(order, *Bm.shape, b.astype(np.float64), ldb)
    Aop = Aop if ta == 111 else Aop.T
    Bop = Bop if tb == 111 else Bop.T
    bound = 4 * (k + 2) * U32 * (abs(alpha) * (np.abs(Aop) @ np.abs(Bop)) + abs(beta) * np.abs(Cm))
    assert np.all(np.abs(got - ref) <= bound)


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
def test_ssyrk_matches_f32_oracle(zh, order, uplo, trans):
    n, k = 150, 90
    rng = np.random.default_rng(n + uplo + trans)
    Am = rng.uniform(-1, 1, (n, k)) if trans == 111 else rng.uniform(-1, 1, (k, n))
    lda = Am.shape[0] if order == COL else Am.shape[1]
    a = _f32(store(order, Am, lda))
    C = rng.uniform(-1, 1, (n, n))
    c = _f32(store(order, C, n))
    c_ref = c.copy()
    assert O.ssyrk(order, uplo, trans, n, k, -1.0, a, lda, 1.0, c_ref, n) == 0
    zh.cblas_ssyrk(order, uplo, trans, n, k, -1.0, P(a), lda, 1.0, P(c), n)
    got, ref = dense(order, n, n, c.astype(np.float64), n), dense(order, n, n, c_ref.astype(np.float64), n)
    tri = np.triu(np.ones((n, n), bool)) if uplo == 121 else np.tril(np.ones((n, n), bool))
    assert np.all(np.abs(got - ref)[tri] <= 4 * (k + 2) * U32 * (k + 1))
    assert np.array_equal(got[~tri], ref[~tri])   # the other triangle is untouched


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("diag", [131, 132])
def test_strsm_matches_f32_oracle(zh, order, side, uplo, trans, diag):
    m, n = 97, 70
    na = m if side == 141 else n
    rng = np.random.default_rng(m + side + uplo + trans + diag + order)
    A = rng.uniform(-1, 1, (na, na)) + 6 * np.eye(na)
    a = _f32(store(order, A, na))
    B = rng.uniform(-1, 1, (m, n))
    ldb = m if order == COL else n
    b = _f32(store(order, B, ldb))
    b_ref = b.copy()
    O.set_trsm_faithful_lda(False)
    try:
        assert O.strsm(order, side, uplo, trans, diag, m, n, 2.0, a, na, b_ref, ldb) == 0
    finally:
        O.set_trsm_faithful_lda(True)
    zh.cblas_strsm(order, side, uplo, trans, diag, m, n, 2.0, P(a), na, P(b), ldb)
    got, ref = b.astype(np.float64), b_ref.astype(np.float64)
    assert np.all(np.abs(got - ref) <= 2e-5 * (np.abs(ref).max() + 1.0)), np.abs(got - ref).max()


def test_f32_device_entries(zh, torch_cuda):
    torch = torch_cuda
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    m, n, k = 200, 130, 64
    rng = np.random.default_rng(9)
    A, B = _f32(rng.uniform(-1, 1, (m, k))), _f32(rng.uniform(-1, 1, (k, n)))
    a = torch.from_numpy(A.T.copy()).cuda()
    b = torch.from_numpy(B.T.copy()).cuda()
    c = torch.zeros(n, m, dtype=torch.float32, device="cuda")
    assert zh.zmlhip_sgemm(COL, 111, 111, m, n, k, 1.0, P(a), m, P(b), k, 0.0, P(c), m, s) == 0
    torch.cuda.synchronize()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(c.cpu().numpy().T - ref).max() <= 4 * (k + 2) * U32 * k
    assert zh.zmlhip_sgemm(COL, 111, 111, m, n, k, 1.0, P(a), m - 1, P(b), k, 0.0, P(c), m, s) == -9


# ---- lange ----------------------------------------------------------------------
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("m,n", [(1, 1), (33, 700), (1000, 257), (4097, 3)])
def test_dlange_matches_numpy(zh, order, m, n):
    rng = np.random.default_rng(m * n + order)
    A = rng.uniform(-5, 5, (m, n))
    ld = (m if order == COL else n) + 2
    a = store(order, A, ld, fill=np.nan)    # the padding must not be read
    assert zh.LAPACKE_dlange(order, b"M", m, n, P(a), ld) == np.abs(A).max()
    assert zh.LAPACKE_dlange(order, b"I", m, n, P(a), ld) == pytest.approx(np.linalg.norm(A, np.inf), rel=1e-14)
    for c in (b"O", b"1"):
        assert zh.LAPACKE_dlange(order, c, m, n, P(a), ld) == pytest.approx(np.linalg.norm(A, 1), rel=1e-14)
    for c in (b"F", b"E"):
        assert zh.LAPACKE_dlange(order, c, m, n, P(a), ld) == pytest.approx(np.linalg.norm(A, "fro"), rel=1e-13)
    assert zh.LAPACKE_dlange(order, b"I", m, n, P(a), ld) == pytest.approx(O.lange(order, "I", m, n, a, ld),
                                                                          rel=1e-14)


def test_dlange_nan_zero_and_invalid(zh, torch_cuda):
    A = np.ones((50, 40))
    A[17, 3] = np.nan
    a = store(COL, A, 50)
    for c in (b"M", b"O", b"I", b"F"):
        assert np.isnan(zh.LAPACKE_dlange(COL, c, 50, 40, P(a), 50))
    Z = np.zeros(6)
    assert zh.LAPACKE_dlange(COL, b"F", 2, 3, P(Z), 2) == 0.0
    assert zh.LAPACKE_dlange(COL, b"M", 0, 3, P(Z), 1) == 0.0
    assert zh.LAPACKE_dlange(5, b"M", 2, 3, P(Z), 2) == -1.0
    assert zh.LAPACKE_dlange(COL, b"Q", 2, 3, P(Z), 2) == -2.0
    # huge entries: the Frobenius norm is scaled, so it does not overflow
    H = np.full((4, 4), 1e300)
    assert zh.LAPACKE_dlange(COL, b"F", 4, 4, P(store(COL, H, 4)), 4) == pytest.approx(4e300, rel=1e-14)
    torch = torch_cuda
    d = torch.from_numpy(store(ROW, np.arange(12.0).reshape(3, 4) - 6, 4)).cuda()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert zh.zmlhip_dlange(ROW, b"I", 3, 4, P(d), 4, P(out), s) == 0
    torch.cuda.synchronize()
    assert out.item() == np.linalg.norm(np.arange(12.0).reshape(3, 4) - 6, np.inf)


# ---- IEEE square root in the Cholesky diagonal kernel ------------------------------
@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_diagonal_spd_potrf_is_exact_sqrt(zh, order, uplo):
    """potrf2.zig:37-53 takes a correctly rounded sqrt (float/sqrt.zig:31-41):
    L of a diagonal SPD matrix must be numpy's (IEEE) sqrt bit for bit, over
    mantissas and exponents that exercise the rounding."""
    n = 700
    rng = np.random.default_rng(11)
    d = rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-600, 600, n)
    d[:4] = [2.0, 3.0, 1.0 + 2.0 ** -52, np.nextafter(4.0, 0)]
    a = store(order, np.diag(d), n)
    assert zh.LAPACKE_dpotrf(order, uplo, n, P(a), n) == 0
    L = dense(order, n, n, a, n)
    assert np.array_equal(np.diag(L).view(np.uint64), np.sqrt(d).view(np.uint64))
    for fn in (zh.LAPACKE_dpotrf2,):
        b = store(order, np.diag(d), n)
        assert fn(order, uplo, n, P(b), n) == 0
        assert np.array_equal(np.diag(dense(order, n, n, b, n)), np.sqrt(d))


def test_diagonal_spd_spotrf_is_exact_sqrt(zh):
    n = 300
    d = (np.random.default_rng(12).uniform(1.0, 2.0, n) * 2.0 ** np.arange(-60, 60, 0.4)[:n].round()).astype(np.float32)
    a = store(COL, np.diag(d.astype(np.float64)), n).astype(np.float32)
    assert zh.LAPACKE_spotrf(COL, b"L", n, P(a), n) == 0
    assert np.array_equal(np.diag(a.reshape(n, n).T), np.sqrt(d))
