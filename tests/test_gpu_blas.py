"""GPU parity of the Level-1/2/3 BLAS through the C ABI (libzmlhip.so).

* the reference's own golden vectors (test/linalg/blas/*.zig), replayed
  through the host-pointer cblas_* drop-ins: exact for gemm/syrk (integer
  data), rel 1e-7 for trsm, exact for iamax/scal/swap;
* seeded random operands against the CPU oracle, every transpose/side/uplo/
  diag/layout combination, with the elementwise bound
  |C - C_ref| <= 2 (k + 2) u (|alpha| |A||B| + |beta| |C|)   (u = 2^-53)
  since the matrix cores fuse and reorder what zml rounds separately.
"""
import numpy as np
import pytest

import oracle as O
from tests import fixture_replay
from tests.helpers import COL, ROW, U64, P, dense, store

pytestmark = pytest.mark.gpu


class CblasBackend:
    def __init__(self, zh):
        self.zh = zh

    def gemm(self, order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        self.zh.cblas_dgemm(order, ta, tb, m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc)

    def trsm(self, order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
        self.zh.cblas_dtrsm(order, side, uplo, ta, diag, m, n, alpha, P(a), lda, P(b), ldb)

    def syrk(self, order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
        self.zh.cblas_dsyrk(order, uplo, trans, n, k, alpha, P(a), lda, beta, P(c), ldc)


@pytest.mark.parametrize("name", ["gemm", "trsm", "syrk"])
def test_reference_golden_vectors_through_cblas(zh, name):
    fails = fixture_replay.replay(name, CblasBackend(zh))
    assert not fails, "\n".join(fails[:10])


def test_reference_iamax_scal_swap_through_cblas(zh):
    fx = fixture_replay.load()
    n = fx["iamax"]["n"]
    x = np.zeros(n)
    x[fx["iamax"]["ones_at"]] = 1.0
    for case in fx["iamax"]["cases"]:
        assert zh.cblas_idamax(case["n"], P(x), case["incx"]) == case["expect"]
    assert zh.cblas_idamax(0, P(x), 1) == 0       # invalid -> 0 (blas.zig:4573-4578)
    i = np.arange(n, dtype=np.float64)
    x1 = i + 1
    zh.cblas_dscal(n, 2.0, P(x1), 1)
    assert np.array_equal(x1, 2 * (i + 1))
    x3 = i + 1
    zh.cblas_dscal(n // 2, 2.0, P(x3), 2)
    assert np.array_equal(x3, np.where(i % 2 == 0, 2 * (i + 1), i + 1))
    x1, x2 = i + 1, i + 2
    zh.cblas_dswap(n, P(x1), 1, P(x2), 1)
    assert np.array_equal(x1, i + 2) and np.array_equal(x2, i + 1)
    x1, x3 = i + 1, n - i + 1
    zh.cblas_dswap(n, P(x1), 1, P(x3), -1)
    assert np.array_equal(x1, i + 2) and np.array_equal(x3, n - i)
    x1, x4 = i + 1, i + 2
    zh.cblas_dswap(n // 2, P(x1), 2, P(x4), 2)
    assert np.array_equal(x1, np.where(i % 2 == 0, i + 2, i + 1))
    assert np.array_equal(x4, np.where(i % 2 == 0, i + 1, i + 2))


def test_idamax_edge_cases(zh):
    x = np.array([np.nan, 5.0, -7.0])
    assert zh.cblas_idamax(3, P(x), 1) == 0            # NaN first element is never passed
    x = np.array([1.0, np.nan, -7.0, 7.0])
    assert zh.cblas_idamax(4, P(x), 1) == 2            # NaN skipped, first of the ties
    x = np.array([1.0, -np.inf, np.inf])
    assert zh.cblas_idamax(3, P(x), 1) == 1
    big = np.random.default_rng(1).uniform(-1, 1, 300001)
    big[123456] = 3.0
    big[200001] = -3.0
    assert zh.cblas_idamax(big.size, P(big), 1) == 123456
    assert zh.cblas_idamax(1, P(big), 1) == 0


def _gemm_bound(alpha, A, B, beta, C, k):
    return 2 * (k + 2) * U64 * (abs(alpha) * (np.abs(A) @ np.abs(B)) + abs(beta) * np.abs(C)) + 1e-300


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("ta", [111, 112])
@pytest.mark.parametrize("tb", [111, 112])
@pytest.mark.parametrize("m,n,k,alpha,beta", [(1, 1, 1, 1.0, 0.0), (37, 53, 29, -1.0, 1.0),
                                              (130, 257, 70, 0.5, -2.0), (256, 128, 300, 1.0, 0.0),
                                              (5, 7, 0, 2.0, 3.0), (64, 64, 64, 0.0, 2.0)])
def test_dgemm_matches_oracle(zh, order, ta, tb, m, n, k, alpha, beta):
    rng = np.random.default_rng(m * 1000 + n + k)
    Am = rng.uniform(-1, 1, (m, k)) if ta == 111 else rng.uniform(-1, 1, (k, m))
    Bm = rng.uniform(-1, 1, (k, n)) if tb == 111 else rng.uniform(-1, 1, (n, k))
    Cm = rng.uniform(-1, 1, (m, n))
    pad = 3
    lda = (Am.shape[0] if order == COL else Am.shape[1]) + pad
    ldb = (Bm.shape[0] if order == COL else Bm.shape[1]) + pad
    ldc = (m if order == COL else n) + pad
    a, b = store(order, Am, lda), store(order, Bm, ldb)
    c = store(order, Cm, ldc, fill=np.nan)
    c_ref = c.copy()
    assert O.gemm(order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c_ref, ldc) == 0
    zh.cblas_dgemm(order, ta, tb, m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc)
    got, ref = dense(order, m, n, c, ldc), dense(order, m, n, c_ref, ldc)
    Aop = Am if ta == 111 else Am.T
    Bop = Bm if tb == 111 else Bm.T
    assert np.all(np.abs(got - ref) <= 2 * _gemm_bound(alpha, Aop, Bop, beta, Cm, max(k, 1)))
    # padding between columns/rows is never written
    mask = np.isnan(c_ref)
    assert np.array_equal(np.isnan(c), mask)


def test_dgemm_beta_zero_ignores_nan_in_c(zh):
    m = n = k = 40
    rng = np.random.default_rng(5)
    A, B = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n))
    a, b = store(COL, A, m), store(COL, B, k)
    c = np.full(m * n, np.nan)
    zh.cblas_dgemm(COL, 111, 111, m, n, k, 1.0, P(a), m, P(b), k, 0.0, P(c), m)
    assert np.all(np.isfinite(c))
    assert np.allclose(dense(COL, m, n, c, m), A @ B, rtol=0, atol=1e-13)


SIDES, UPLOS, TRANS, DIAGS = [141, 142], [121, 122], [111, 112], [131, 132]


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("side", SIDES)
@pytest.mark.parametrize("uplo", UPLOS)
@pytest.mark.parametrize("ta", TRANS)
@pytest.mark.parametrize("diag", DIAGS)
def test_dtrsm_matches_oracle(zh, order, side, uplo, ta, diag):
    m, n, alpha = 77, 45, 1.5
    na = m if side == 141 else n
    rng = np.random.default_rng(side + uplo + ta + diag + order)
    T = rng.uniform(-1, 1, (na, na)) + np.eye(na) * (na / 2.0)   # well conditioned
    Bm = rng.uniform(-1, 1, (m, n))
    lda = na + 2
    ldb = (m if order == COL else n) + 2
    a = store(order, T, lda)
    b = store(order, Bm, ldb)
    b_ref = b.copy()
    O.set_trsm_faithful_lda(False)
    try:
        assert O.trsm(order, side, uplo, ta, diag, m, n, alpha, a, lda, b_ref, ldb) == 0
    finally:
        O.set_trsm_faithful_lda(True)
    zh.cblas_dtrsm(order, side, uplo, ta, diag, m, n, alpha, P(a), lda, P(b), ldb)
    got, ref = dense(order, m, n, b, ldb), dense(order, m, n, b_ref, ldb)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-13 * max(scale, 1.0)


def test_dtrsm_large_blocked(zh):
    m, n = 700, 300
    rng = np.random.default_rng(9)
    L = np.tril(rng.uniform(-1, 1, (m, m))) + np.eye(m) * m
    X = rng.uniform(-1, 1, (m, n))
    B = L @ X
    a, b = store(COL, L, m), store(COL, B, m)
    zh.cblas_dtrsm(COL, 141, 122, 111, 131, m, n, 1.0, P(a), m, P(b), m)
    assert np.abs(dense(COL, m, n, b, m) - X).max() < 1e-12


def test_dtrsm_alpha_zero_zeroes_b(zh):
    m, n = 9, 4
    a = np.eye(m).ravel()
    b = np.full(m * n, np.nan)
    zh.cblas_dtrsm(COL, 141, 122, 111, 131, m, n, 0.0, P(a), m, P(b), m)
    assert np.array_equal(b, np.zeros(m * n))


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("uplo", UPLOS)
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("n,k,alpha,beta", [(5, 3, 2.0, 3.0), (150, 70, -1.0, 1.0), (260, 33, 0.7, 0.0),
                                            (2100, 40, -1.0, 1.0)])
def test_dsyrk_matches_oracle(zh, order, uplo, trans, n, k, alpha, beta):
    rng = np.random.default_rng(n + k + uplo + trans)
    Am = rng.uniform(-1, 1, (n, k)) if trans == 111 else rng.uniform(-1, 1, (k, n))
    Cm = rng.uniform(-1, 1, (n, n))
    lda = (Am.shape[0] if order == COL else Am.shape[1]) + 1
    ldc = n + 1
    a = store(order, Am, lda)
    c = store(order, Cm, ldc)
    c_ref = c.copy()
    assert O.syrk(order, uplo, trans, n, k, alpha, a, lda, beta, c_ref, ldc) == 0
    zh.cblas_dsyrk(order, uplo, trans, n, k, alpha, P(a), lda, beta, P(c), ldc)
    got, ref = dense(order, n, n, c, ldc), dense(order, n, n, c_ref, ldc)
    Aop = Am if trans == 111 else Am.T
    bound = 2 * _gemm_bound(alpha, Aop, Aop.T, beta, Cm, k)
    assert np.all(np.abs(got - ref) <= bound)
    # the other triangle is untouched
    other = np.triu(np.ones((n, n)), 1) if uplo == 122 else np.tril(np.ones((n, n)), -1)
    assert np.array_equal(got[other == 1], Cm[other == 1])


@pytest.mark.parametrize("order", [COL, ROW])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_dgemv_matches_oracle(zh, order, trans, incx, incy):
    m, n = 301, 177
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (m, n))
    lda = (m if order == COL else n) + 2
    a = store(order, A, lda)
    lenx, leny = (n, m) if trans == 111 else (m, n)
    x = rng.uniform(-1, 1, lenx * abs(incx))
    y = rng.uniform(-1, 1, leny * abs(incy))
    y_ref = y.copy()
    assert O.gemv(order, trans, m, n, 0.75, a, lda, x, incx, -0.5, y_ref, incy) == 0
    zh.cblas_dgemv(order, trans, m, n, 0.75, P(a), lda, P(x), incx, -0.5, P(y), incy)
    assert np.abs(y - y_ref).max() < 1e-12
