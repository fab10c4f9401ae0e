"""Pins the CPU oracle (oracle/zml_oracle.c) before anything is compared to it.

* the reference's own golden vectors (test/linalg/blas/*.zig) replayed through
  the oracle's gemm/trsm/syrk/iamax/scal/swap;
* self-consistency of zml's blocked/recursive drivers against their closed
  forms (bitwise), which is what lets a GPU result be checked against the
  unblocked LU at large n;
* mathematics that a plausible mistake would break: hand-solved 3x3 LU,
  identity / permutation / singular / diagonal special cases, ||PA - LU||,
  solutions against numpy's LAPACK, the splitmix64 published constants;
* planted mistakes (in the oracle's inputs/outputs, on the CPU) that these
  checks must catch.
"""
import numpy as np
import pytest

import oracle as O
from tests import fixture_replay


class OracleBackend:
    gemm = staticmethod(O.gemm)
    trsm = staticmethod(O.trsm)
    syrk = staticmethod(O.syrk)


@pytest.mark.parametrize("name", ["gemm", "trsm", "syrk"])
def test_reference_golden_vectors(oracle_lib, name):
    fails = fixture_replay.replay(name, OracleBackend)
    assert not fails, "\n".join(fails[:10])


def test_reference_iamax_scal_swap(oracle_lib):
    fx = fixture_replay.load()
    n = fx["iamax"]["n"]
    x = np.zeros(n)
    x[fx["iamax"]["ones_at"]] = 1.0
    for case in fx["iamax"]["cases"]:
        assert O.iamax(case["n"], x, case["incx"]) == case["expect"]
    assert O.iamax(0, x, 1) == -1 and O.iamax(5, x, 0) == -1  # iamax.zig:19

    i = np.arange(n, dtype=np.float64)
    x1 = i + 1
    O.scal(n, 2.0, x1, 1)
    assert np.array_equal(x1, 2 * (i + 1))
    x3 = i + 1
    O.scal(n // 2, 2.0, x3, 2)
    assert np.array_equal(x3, np.where(i % 2 == 0, 2 * (i + 1), i + 1))

    x1, x2 = i + 1, i + 2
    O.swap(n, x1, 1, x2, 1)
    assert np.array_equal(x1, i + 2) and np.array_equal(x2, i + 1)
    x1, x3 = i + 1, n - i + 1
    O.swap(n, x1, 1, x3, -1)
    assert np.array_equal(x1, i + 2) and np.array_equal(x3, n - i)
    x1, x4 = i + 1, i + 2
    O.swap(n // 2, x1, 2, x4, 2)
    assert np.array_equal(x1, np.where(i % 2 == 0, i + 2, i + 1))
    assert np.array_equal(x4, np.where(i % 2 == 0, i + 1, i + 2))


def test_planted_gemm_mistake_is_caught(oracle_lib):
    """A transposed operand in the oracle would fail the golden vectors."""
    class Transposed:
        @staticmethod
        def gemm(order, ta, tb, *rest):
            flip = {111: 112, 112: 111}
            return O.gemm(order, ta, flip.get(tb, tb), *rest)
        trsm = staticmethod(O.trsm)
        syrk = staticmethod(O.syrk)
    try:
        fails = fixture_replay.replay("gemm", Transposed)
    except Exception:
        fails = ["raised"]
    assert fails


# ---------------------------------------------------------------------------
def _mat(order, m, n, seed, lo=-1.0, hi=1.0):
    return O.fill_uniform(order, m, n, seed, lo, hi)


def _to_dense(order, m, n, a, ld):
    if order == O.COL:
        return a.reshape(-1, ld)[:n, :m].T.copy()
    return a.reshape(-1, ld)[:m, :n].copy()


@pytest.mark.parametrize("order", [O.COL, O.ROW])
@pytest.mark.parametrize("m,n", [(5, 5), (64, 64), (65, 65), (200, 200), (300, 170), (130, 260), (257, 257)])
def test_getrf_matches_unblocked_closed_form_bitwise(oracle_lib, order, m, n):
    a = _mat(order, m, n, 1000 + m + n)
    b = a.copy()
    ipa = np.zeros(min(m, n), np.int32)
    ipb = np.zeros(min(m, n), np.int32)
    ia = O.getrf(order, m, n, a, m if order == O.COL else n, ipa)
    ib = O.getrf_unblocked(order, m, n, b, m if order == O.COL else n, ipb)
    assert ia == ib == 0
    assert np.array_equal(ipa, ipb)
    assert np.array_equal(a, b)  # bitwise, including rounding of every update


@pytest.mark.parametrize("m,n", [(200, 200), (333, 333), (400, 250), (250, 400)])
def test_getrf_fast_matches_unblocked_bitwise(oracle_lib, m, n):
    a = _mat(O.COL, m, n, 7 + m)
    b = a.copy()
    ipa = np.zeros(min(m, n), np.int32)
    ipb = np.zeros(min(m, n), np.int32)
    assert O.getrf_fast(m, n, a, m, ipa, 4) == O.getrf_unblocked(O.COL, m, n, b, m, ipb)
    assert np.array_equal(ipa, ipb)
    assert np.array_equal(a, b)


def test_row_and_col_major_getrf_agree_bitwise(oracle_lib):
    n = 150
    ar = _mat(O.ROW, n, n, 5)
    ac = _mat(O.COL, n, n, 5)
    assert np.array_equal(_to_dense(O.ROW, n, n, ar, n), _to_dense(O.COL, n, n, ac, n))
    ipr = np.zeros(n, np.int32)
    ipc = np.zeros(n, np.int32)
    O.getrf(O.ROW, n, n, ar, n, ipr)
    O.getrf(O.COL, n, n, ac, n, ipc)
    assert np.array_equal(ipr, ipc)
    assert np.array_equal(_to_dense(O.ROW, n, n, ar, n), _to_dense(O.COL, n, n, ac, n))


def _perm_from_ipiv(ipiv, m):
    p = np.arange(m)
    for i, ip in enumerate(ipiv):
        p[[i, ip - 1]] = p[[ip - 1, i]]
    return p


@pytest.mark.parametrize("order", [O.COL, O.ROW])
def test_getrf_reconstructs_pa(oracle_lib, order):
    n = 180
    a = _mat(order, n, n, 11)
    A = _to_dense(order, n, n, a, n)
    ip = np.zeros(n, np.int32)
    assert O.getrf(order, n, n, a, n, ip) == 0
    LU = _to_dense(order, n, n, a, n)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    P = _perm_from_ipiv(ip, n)
    err = np.abs(A[P] - L @ U).max() / np.abs(A).max()
    assert err < 50 * n * 2.0 ** -53
    assert np.all(ip >= np.arange(1, n + 1)) and np.all(ip <= n)
    # first-max pivoting bounds every multiplier by 1
    assert np.abs(np.tril(LU, -1)).max() <= 1.0


def test_getrf_hand_solved_3x3(oracle_lib):
    """A = [[1,2,3],[4,5,6],[7,8,10]] worked by hand with zml's rules:
    step 1 pivots on 7 (row 3), step 2 on 2 - 8/7 (row 3 again)."""
    A = np.array([[1.0, 2, 3], [4, 5, 6], [7, 8, 10]])
    a = A.copy().ravel()  # row-major
    ip = np.zeros(3, np.int32)
    assert O.getrf(O.ROW, 3, 3, a, 3, ip) == 0
    assert list(ip) == [3, 3, 3]
    r = 1.0 / 7.0
    l_a, l_b = 1.0 * r, 4.0 * r            # multipliers of rows [1,2,3] and [4,5,6]
    u22, u23 = 2.0 - l_a * 8.0, 3.0 - l_a * 10.0
    a32, a33 = 5.0 - l_b * 8.0, 6.0 - l_b * 10.0
    assert abs(u22) > abs(a32)
    l32 = a32 * (1.0 / u22)
    u33 = a33 - l32 * u23
    expect = np.array([[7.0, 8, 10], [l_a, u22, u23], [l_b, l32, u33]])
    assert np.array_equal(a.reshape(3, 3), expect)


def test_getrf_special_cases(oracle_lib):
    n = 70
    eye = np.eye(n).ravel()
    ip = np.zeros(n, np.int32)
    assert O.getrf(O.COL, n, n, eye, n, ip) == 0
    assert np.array_equal(ip, np.arange(1, n + 1))
    assert np.array_equal(eye, np.eye(n).ravel())
    # permutation matrix: exact factors, L = I, U = I
    rng = np.random.default_rng(3)
    perm = rng.permutation(n)
    P = np.eye(n)[perm]
    a = P.T.copy().ravel()  # column-major storage of P
    ip = np.zeros(n, np.int32)
    assert O.getrf(O.COL, n, n, a, n, ip) == 0
    assert np.array_equal(a, np.eye(n).ravel())
    assert np.array_equal(P[_perm_from_ipiv(ip, n)], np.eye(n))
    # a zero column gives info = first zero column (1-based), factorization continues
    a = _mat(O.COL, n, n, 9)
    A = a.reshape(n, n)
    A[17, :] = 0.0  # column 17 in column-major storage
    ip = np.zeros(n, np.int32)
    assert O.getrf(O.COL, n, n, a, n, ip) == 18
    b = _mat(O.COL, n, n, 9)
    b.reshape(n, n)[17, :] = 0.0
    ip2 = np.zeros(n, np.int32)
    assert O.getrf_unblocked(O.COL, n, n, b, n, ip2) == 18
    assert np.array_equal(ip, ip2) and np.array_equal(a, b)
    # invalid arguments
    assert O.getrf(O.COL, -1, 3, a, 3, ip) == -1
    assert O.getrf(O.COL, 5, 5, a, 4, ip) == -1


@pytest.mark.parametrize("order", [O.COL, O.ROW])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_gesv_solves(oracle_lib, order, nrhs):
    n = 160
    a = _mat(order, n, n, 21)
    A = _to_dense(order, n, n, a, n)
    X = _to_dense(order, n, nrhs, _mat(order, n, nrhs, 22), n if order == O.COL else nrhs)
    B = A @ X
    b = (B.T if order == O.COL else B).copy().ravel()
    ip = np.zeros(n, np.int32)
    assert O.gesv(order, n, nrhs, a, n, ip, b, n if order == O.COL else nrhs) == 0
    Xs = _to_dense(order, n, nrhs, b, n if order == O.COL else nrhs)
    assert np.abs(Xs - np.linalg.solve(A, B)).max() < 1e-8
    assert np.abs(A @ Xs - B).max() / (np.abs(A).sum(1).max() * np.abs(Xs).max()) < 16 * n * 2.0 ** -53
    # transposed solve with the same factors
    bt = (A.T @ X)
    bt = (bt.T if order == O.COL else bt).copy().ravel()
    O.getrs(order, O.TRANS, n, nrhs, a, n, ip, bt, n if order == O.COL else nrhs)
    assert np.abs(_to_dense(order, n, nrhs, bt, n if order == O.COL else nrhs) - X).max() < 1e-8


def _spd(n, seed, mu=1.0):
    B = O.fill_uniform(O.ROW, n, n, seed).reshape(n, n)
    return B @ B.T / n + mu * np.eye(n)


@pytest.mark.parametrize("n", [5, 64, 100, 200])
def test_potrf_lower_col_major_matches_closed_form_bitwise(oracle_lib, n):
    A = _spd(n, 31 + n)
    a = A.T.copy().ravel()
    b = a.copy()
    assert O.potrf(O.COL, O.LOWER, n, a, n) == 0
    assert O.potrf_lower_unblocked(n, b, n) == 0
    La = np.tril(a.reshape(n, n).T)
    Lb = np.tril(b.reshape(n, n).T)
    assert np.array_equal(La, Lb)


@pytest.mark.parametrize("order", [O.COL, O.ROW])
@pytest.mark.parametrize("uplo", [O.LOWER, O.UPPER])
def test_potrf_potrs_math(oracle_lib, order, uplo):
    n, nrhs = 150, 4
    A = _spd(n, 41)
    a = (A.T if order == O.COL else A).copy().ravel()
    assert O.potrf(order, uplo, n, a, n) == 0
    F = _to_dense(order, n, n, a, n)
    if uplo == O.LOWER:
        L = np.tril(F)
        assert np.abs(L @ L.T - A).max() < 1e-12 * n
    else:
        U = np.triu(F)
        assert np.abs(U.T @ U - A).max() < 1e-12 * n
    X = np.arange(n * nrhs, dtype=np.float64).reshape(n, nrhs) / (n * nrhs) - 0.5
    B = A @ X
    # potrs checks ldb >= n in both layouts (potrs.zig:30); use ldb = n
    if order == O.COL:
        b = B.T.copy().ravel()
        ldb = n
    else:
        b = np.zeros((n, n))
        b[:, :nrhs] = B
        b = b.ravel()
        ldb = n
    assert O.potrs(order, uplo, n, nrhs, a, n, b, ldb) == 0
    Xs = b.reshape(nrhs, n).T if order == O.COL else b.reshape(n, n)[:, :nrhs]
    assert np.abs(Xs - X).max() < 1e-11


def test_potrs_row_major_small_ldb_rejected(oracle_lib):
    """zml's potrs checks ldb >= n even for row-major (potrs.zig:30)."""
    n = 10
    a = np.eye(n).ravel()
    b = np.ones(n * 2)
    assert O.potrs(O.ROW, O.LOWER, n, 2, a, n, b, 2) == -1


def test_potrf_special_cases(oracle_lib):
    n = 90
    d = np.arange(1, n + 1, dtype=np.float64)
    a = np.diag(d).ravel()
    assert O.potrf(O.COL, O.LOWER, n, a, n) == 0
    assert np.array_equal(np.diag(a.reshape(n, n)), np.sqrt(d))
    A = _spd(n, 2)
    A[70, 70] = -1.0
    a = A.T.copy().ravel()
    assert O.potrf(O.COL, O.LOWER, n, a, n) == 71
    A = _spd(n, 2)
    A[3, 3] = np.nan
    a = A.T.copy().ravel()
    assert O.potrf(O.COL, O.LOWER, n, a, n) == 4


def test_pivots_agree_with_lapack_on_separated_matrix(oracle_lib):
    """scipy's getrf uses the same partial pivoting rule with another op order;
    on a random matrix pivots agree wherever the pivot gap is not tiny."""
    scipy_linalg = pytest.importorskip("scipy.linalg")
    n = 300
    a = _mat(O.COL, n, n, 77)
    A = _to_dense(O.COL, n, n, a, n)
    ip = np.zeros(n, np.int32)
    O.getrf(O.COL, n, n, a, n, ip)
    lu, piv = scipy_linalg.lu_factor(A)
    assert O.min_pivot_gap(O.COL, n, _mat(O.COL, n, n, 77), n) > 1e-9
    assert np.array_equal(piv + 1, ip)
    assert np.abs(lu - _to_dense(O.COL, n, n, a, n)).max() < 1e-11


def _splitmix_ref(seed, ctr):
    M = (1 << 64) - 1
    z = (seed * 0xD1B54A32D192ED03 + (ctr + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_rng_is_splitmix64(oracle_lib):
    # splitmix64 from state 0: first output is 0xE220A8397B1DCDAF (published).
    assert _splitmix_ref(0, 0) == 0xE220A8397B1DCDAF
    for seed, ctr in [(0, 0), (42, 0), (42, 123456789), (2**63, 2**40)]:
        assert O.rng_u64(seed, ctr) == _splitmix_ref(seed, ctr)
    u = O.fill_uniform(O.ROW, 3, 4, 42)
    ref = np.array([((_splitmix_ref(42, c) >> 11) * 2.0 ** -53) * 2 - 1 for c in range(12)])
    assert np.array_equal(u, ref)


def test_backward_error_metric(oracle_lib):
    n = 50
    A = _to_dense(O.ROW, n, n, _mat(O.ROW, n, n, 3), n)
    x = np.linspace(-1, 1, n)
    b = A @ x
    a = A.ravel().copy()
    be = O.backward_error(O.ROW, n, a, n, x, b)
    assert be < 16
    x2 = x.copy()
    x2[0] += 1e-6
    assert O.backward_error(O.ROW, n, a, n, x2, b) > 1e3
