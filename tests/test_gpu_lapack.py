"""GPU parity of the LAPACK drivers through the C ABI (libzmlhip.so).

The bar for getrf is the one zml's own CPU getrf sets: identical pivot
vectors (checked against the oracle, which is bitwise zml's getrf), the
first 64-column leaf bitwise equal (same operation sequence, no FMA), the
factors within a rounding-error bound, solutions within 1e-10 relative
forward error of the oracle's, and the normwise backward error
||b - Ax||_inf / (||A||_inf ||x||_inf n u) <= 16 with u = 2^-53.
"""
import os

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, ROW, U64, P, dense, store

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _lu_dense(order, m, n, buf, ld):
    LU = dense(order, m, n, buf, ld)
    k = min(m, n)
    L = np.tril(LU[:, :k], -1) + np.eye(m, k)
    U = np.triu(LU[:k, :])
    return L, U


def _perm(ipiv, m):
    p = np.arange(m)
    for i, ip in enumerate(ipiv):
        p[[i, ip - 1]] = p[[ip - 1, i]]
    return p


def _backward_error(A, x, b):
    r = b - A @ x
    return np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(x).max() * A.shape[0] * U64)


@pytest.mark.parametrize("order", [ROW, COL])
def test_config1_lu_n1024_pivots_and_solution(zh, lu_panel, order):
    """Config 1: n = 1024, uniform[-1, 1] (seed 42), x_true uniform (seed 43), b = A x_true."""
    n = 1024
    a = O.fill_uniform(order, n, n, 42)
    A = dense(order, n, n, a, n)
    x_true = O.fill_uniform(ROW, n, 1, 43)
    b = A @ x_true
    # oracle
    a_ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    assert O.getrf(order, n, n, a_ref, n, ip_ref) == 0
    x_ref = b.copy()
    O.getrs(order, 111, n, 1, a_ref, n, ip_ref, x_ref, n if order == COL else 1)
    # GPU
    a_gpu = a.copy()
    ip = np.zeros(n, np.int32)
    assert zh.LAPACKE_dgetrf(order, n, n, P(a_gpu), n, P(ip)) == 0
    assert np.array_equal(ip, ip_ref), f"first pivot mismatch at {np.argmax(ip != ip_ref)}"
    LUg, LUr = dense(order, n, n, a_gpu, n), dense(order, n, n, a_ref, n)
    assert np.array_equal(LUg[:, :64], LUr[:, :64])            # first leaf: bitwise
    assert np.abs(LUg - LUr).max() < 1e-11 * np.abs(LUr).max()
    L, U = _lu_dense(order, n, n, a_gpu, n)
    assert np.abs(A[_perm(ip, n)] - L @ U).max() < 64 * n * U64
    x = b.copy()
    assert zh.LAPACKE_dgetrs(order, b"N", n, 1, P(a_gpu), n, P(ip), P(x), n if order == COL else 1) == 0
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-10
    assert _backward_error(A, x, b) <= 16


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("m,n", [(1, 1), (3, 3), (64, 64), (65, 65), (300, 300), (513, 513),
                                 (700, 260), (260, 700), (1000, 999)])
def test_dgetrf_shapes_match_oracle(zh, lu_panel, order, m, n):
    a = O.fill_uniform(order, m, n, 1000 + m * 7 + n)
    ld = (m if order == COL else n) + 5
    A = dense(order, m, n, a, m if order == COL else n)
    buf = store(order, A, ld)
    ref = buf.copy()
    ip_ref = np.zeros(min(m, n), np.int32)
    info_ref = O.getrf(order, m, n, ref, ld, ip_ref)
    ip = np.zeros(min(m, n), np.int32)
    info = zh.LAPACKE_dgetrf(order, m, n, P(buf), ld, P(ip))
    assert info == info_ref == 0
    assert np.array_equal(ip, ip_ref)
    got, exp = dense(order, m, n, buf, ld), dense(order, m, n, ref, ld)
    assert np.abs(got - exp).max() <= 1e-11 * max(1.0, np.abs(exp).max())


def test_dgetrf_singular_and_special(zh, lu_panel):
    n = 200
    # identity: no interchanges, factors unchanged
    eye = np.eye(n).ravel()
    ip = np.zeros(n, np.int32)
    assert zh.LAPACKE_dgetrf(COL, n, n, P(eye), n, P(ip)) == 0
    assert np.array_equal(ip, np.arange(1, n + 1)) and np.array_equal(eye, np.eye(n).ravel())
    # exactly singular: zero column 150 -> info 151, same pivots as zml
    a = O.fill_uniform(COL, n, n, 3)
    a.reshape(n, n)[150, :] = 0.0
    ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    info_ref = O.getrf(COL, n, n, ref, n, ip_ref)
    info = zh.LAPACKE_dgetrf(COL, n, n, P(a), n, P(ip))
    assert info == info_ref == 151
    assert np.array_equal(ip, ip_ref)
    # permutation matrix: exact
    perm = np.random.default_rng(4).permutation(n)
    Pm = np.eye(n)[perm]
    a = Pm.ravel().copy()   # row-major
    assert zh.LAPACKE_dgetrf(ROW, n, n, P(a), n, P(ip)) == 0
    assert np.array_equal(a, np.eye(n).ravel())
    # invalid arguments: LAPACKE numbering
    assert zh.LAPACKE_dgetrf(COL, -1, n, P(a), n, P(ip)) == -2
    assert zh.LAPACKE_dgetrf(COL, n, n, P(a), n - 1, P(ip)) == -5
    assert zh.LAPACKE_dgetrf(7, n, n, P(a), n, P(ip)) == -1


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("trans", [b"N", b"T"])
@pytest.mark.parametrize("n,nrhs,ldb_extra", [(333, 3, 0), (333, 1, 0), (1100, 1, 2)])
def test_dgetrs_and_dgesv(zh, lu_panel, order, trans, n, nrhs, ldb_extra):
    """nrhs = 1 takes the one-vector path (interchanges in LDS chunks, one
    persistent launch per triangle), also with a padded leading dimension."""
    a = O.fill_uniform(order, n, n, 77)
    A = dense(order, n, n, a, n)
    X = np.random.default_rng(1).uniform(-1, 1, (n, nrhs))
    Aop = A if trans == b"N" else A.T
    B = Aop @ X
    ldb = (n if order == COL else nrhs) + ldb_extra
    ip = np.zeros(n, np.int32)
    lu = a.copy()
    assert zh.LAPACKE_dgetrf(order, n, n, P(lu), n, P(ip)) == 0
    b = store(order, B, ldb)
    assert zh.LAPACKE_dgetrs(order, trans, n, nrhs, P(lu), n, P(ip), P(b), ldb) == 0
    # oracle solution with the oracle's own factors
    lu_ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    O.getrf(order, n, n, lu_ref, n, ip_ref)
    b_ref = store(order, B, ldb)
    O.getrs(order, 111 if trans == b"N" else 112, n, nrhs, lu_ref, n, ip_ref, b_ref, ldb)
    Xg, Xr = dense(order, n, nrhs, b, ldb), dense(order, n, nrhs, b_ref, ldb)
    assert np.abs(Xg - Xr).max() / np.abs(Xr).max() < 1e-10
    if trans == b"N":
        a2 = a.copy()
        b2 = store(order, B, ldb)
        assert zh.LAPACKE_dgesv(order, n, nrhs, P(a2), n, P(ip), P(b2), ldb) == 0
        assert np.abs(dense(order, n, nrhs, b2, ldb) - Xr).max() / np.abs(Xr).max() < 1e-10


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("incx", [1, -1])
def test_dlaswp_matches_oracle(zh, order, incx):
    m, n = 50, 37
    A = np.random.default_rng(2).uniform(-1, 1, (m, n))
    ld = (m if order == COL else n) + 1
    ipiv = np.random.default_rng(3).integers(1, m + 1, size=m).astype(np.int32)
    ipiv = np.maximum(ipiv, np.arange(1, m + 1, dtype=np.int32))
    a = store(order, A, ld)
    ref = a.copy()
    assert O.laswp(order, n, ref, ld, 3, 40, ipiv, incx) == 0
    assert zh.LAPACKE_dlaswp(order, n, P(a), ld, 3, 40, P(ipiv), incx) == 0
    assert np.array_equal(a, ref)


def _spd(n, seed, mu=1.0):
    B = O.fill_uniform(ROW, n, n, seed).reshape(n, n)
    return B @ B.T / n + mu * np.eye(n)


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("n", [1, 7, 128, 129, 500, 1100])
def test_dpotrf_matches_oracle(zh, order, uplo, n):
    A = _spd(n, 40 + n)
    a = store(order, A, n)
    ref = a.copy()
    info_ref = O.potrf(order, 122 if uplo == b"L" else 121, n, ref, n)
    info = zh.LAPACKE_dpotrf(order, uplo, n, P(a), n)
    assert info == info_ref == 0
    got, exp = dense(order, n, n, a, n), dense(order, n, n, ref, n)
    tri = np.tril if uplo == b"L" else np.triu
    assert np.abs(tri(got) - tri(exp)).max() < 1e-12
    # the other triangle is not referenced
    other = np.triu(np.ones((n, n)), 1) if uplo == b"L" else np.tril(np.ones((n, n)), -1)
    assert np.array_equal(got[other == 1], A[other == 1])


def test_dpotrf_not_positive_definite(zh):
    n = 300
    A = _spd(n, 8)
    A[200, 200] = -5.0
    a = store(COL, A, n)
    ref = a.copy()
    assert zh.LAPACKE_dpotrf(COL, b"L", n, P(a), n) == O.potrf(COL, 122, n, ref, n) == 201
    # the factor of the leading minor of order info - 1 agrees with zml's (what
    # lies beyond it depends on how far each blocking got before stopping)
    got, exp = dense(COL, n, n, a, n), dense(COL, n, n, ref, n)
    assert np.abs(np.tril(got)[:200, :200] - np.tril(exp)[:200, :200]).max() < 1e-12
    A = _spd(n, 8)
    A[5, 5] = np.nan
    a = store(COL, A, n)
    assert zh.LAPACKE_dpotrf(COL, b"L", n, P(a), n) == 6


@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_dpotrf_not_positive_definite_later_block_column(zh, uplo):
    """A failing pivot in the third 512-column block column."""
    n = 1300
    A = _spd(n, 9)
    A[1100, 1100] = -5.0
    a = store(COL, A, n)
    ref = a.copy()
    code = 122 if uplo == b"L" else 121
    assert zh.LAPACKE_dpotrf(COL, uplo, n, P(a), n) == O.potrf(COL, code, n, ref, n) == 1101
    got, exp = dense(COL, n, n, a, n), dense(COL, n, n, ref, n)
    tri = np.tril if uplo == b"L" else np.triu
    assert np.abs(tri(got)[:1100, :1100] - tri(exp)[:1100, :1100]).max() < 1e-12


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("n", [400, 1100])
@pytest.mark.parametrize("nrhs", [16, 1])
def test_dpotrs_dposv(zh, order, uplo, n, nrhs):
    A = _spd(n, 12)
    X = np.random.default_rng(5).uniform(-1, 1, (n, nrhs))
    B = A @ X
    ldb = n if order == COL else nrhs
    a = store(order, A, n)
    b = store(order, B, ldb)
    assert zh.LAPACKE_dposv(order, uplo, n, nrhs, P(a), n, P(b), ldb) == 0
    Xg = dense(order, n, nrhs, b, ldb)
    assert np.abs(Xg - X).max() < 1e-11
    b2 = store(order, B, ldb)
    assert zh.LAPACKE_dpotrs(order, uplo, n, nrhs, P(a), n, P(b2), ldb) == 0
    assert np.abs(dense(order, n, nrhs, b2, ldb) - X).max() < 1e-11


def test_lu_n16384_pivots_match_zml(zh, lu_panel):
    """Config 2 size: pivots of the 16384 x 16384 seeded matrix equal the
    golden vector written by tests/golden/make_lu_pivots.py (zml's getrf)."""
    path = os.path.join(GOLDEN, "lu_pivots_n16384_seed42.npz")
    if not os.path.exists(path):
        pytest.fail("golden pivots missing: run tests/golden/make_lu_pivots.py 16384")
    g = np.load(path)
    n = 16384
    a = O.fill_uniform(COL, n, n, 42)
    ip = np.zeros(n, np.int32)
    info = zh.LAPACKE_dgetrf(COL, n, n, P(a), n, P(ip))
    assert info == int(g["info"]) == 0
    mism = np.nonzero(ip != g["ipiv"])[0]
    assert mism.size == 0, f"pivots differ from zml's first at column {mism[0]}"
    diag = np.diagonal(a.reshape(n, n).T)
    assert np.abs(diag - g["diag"]).max() <= 1e-9 * np.abs(g["diag"]).max()


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("trans", [b"N", b"T"])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129])
def test_one_vector_getrs_potrs_small_orders(zh, order, trans, n):
    """The one-vector solves at orders around their 64-row block: partial and
    single blocks, one row, against the oracle's getrs / potrs."""
    a = O.fill_uniform(order, n, n, 90 + n)
    A = dense(order, n, n, a, n)
    x0 = np.random.default_rng(n).uniform(-1, 1, (n, 1))
    B = (A if trans == b"N" else A.T) @ x0
    ldb = n if order == COL else 1
    ip = np.zeros(n, np.int32)
    lu = a.copy()
    assert zh.LAPACKE_dgetrf(order, n, n, P(lu), n, P(ip)) == 0
    b = store(order, B, ldb)
    assert zh.LAPACKE_dgetrs(order, trans, n, 1, P(lu), n, P(ip), P(b), ldb) == 0
    lu_ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    O.getrf(order, n, n, lu_ref, n, ip_ref)
    b_ref = store(order, B, ldb)
    O.getrs(order, 111 if trans == b"N" else 112, n, 1, lu_ref, n, ip_ref, b_ref, ldb)
    Xg, Xr = dense(order, n, 1, b, ldb), dense(order, n, 1, b_ref, ldb)
    assert np.abs(Xg - Xr).max() <= 1e-10 * max(1.0, np.abs(Xr).max())
    if trans == b"N":
        S = _spd(n, 30 + n)
        s_ = store(order, S, n)
        assert zh.LAPACKE_dpotrf(order, b"L", n, P(s_), n) == 0
        y = S @ x0
        yb = store(order, y, ldb)
        assert zh.LAPACKE_dpotrs(order, b"L", n, 1, P(s_), n, P(yb), ldb) == 0
        assert np.abs(dense(order, n, 1, yb, ldb) - x0).max() < 1e-11
