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


# ---- gemv / trsv: the reference's vectors (test/linalg/blas/gemv.zig, trsv.zig)
class OracleBackendL2(OracleBackend):
    gemv = staticmethod(O.gemv)
    trsv = staticmethod(O.trsv)


@pytest.mark.parametrize("name", ["gemv", "trsv"])
def test_reference_level2_golden_vectors(oracle_lib, name):
    fails = fixture_replay.replay(name, OracleBackendL2)
    assert not fails, "\n".join(fails[:10])


def test_planted_trsv_mistake_is_caught(oracle_lib):
    """A trsv that ignored the unit-diagonal flag fails the reference's vectors."""
    class NoUnit(OracleBackendL2):
        @staticmethod
        def trsv(order, uplo, trans, diag, *rest):
            return O.trsv(order, uplo, trans, 131, *rest)
    assert fixture_replay.replay("trsv", NoUnit)


@pytest.mark.parametrize("order", [O.COL, O.ROW])
@pytest.mark.parametrize("uplo", [O.UPPER, O.LOWER])
@pytest.mark.parametrize("trans", [O.NOTRANS, O.TRANS])
@pytest.mark.parametrize("diag", [O.NONUNIT, O.UNIT])
@pytest.mark.parametrize("incx", [1, -3])
def test_trsv_solves_the_triangular_system(oracle_lib, order, uplo, trans, diag, incx):
    """x = op(T)^-1 b against numpy's solve of the dense triangle."""
    n = 37
    rng = np.random.default_rng(n + incx + order + uplo + trans + diag)
    M = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    T = np.triu(M) if uplo == O.UPPER else np.tril(M)
    if diag == O.UNIT:
        np.fill_diagonal(T, 1.0)
    a = (T.T if order == O.COL else T).copy().ravel()
    if diag == O.UNIT:   # the stored diagonal must not be read
        av = a.reshape(n, n)
        np.fill_diagonal(av, 7.0)
    b = rng.uniform(-1, 1, n)
    x = np.zeros(abs(incx) * (n - 1) + 1)
    x[::abs(incx)] = b if incx > 0 else b[::-1]
    assert O.trsv(order, uplo, trans, diag, n, a, n, x, incx) == 0
    got = x[::abs(incx)] if incx > 0 else x[::abs(incx)][::-1]
    want = np.linalg.solve(T if trans == O.NOTRANS else T.T, b)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)
    assert O.trsv(order, uplo, trans, diag, n, a, n, x, 0) == -1   # trsv.zig:76


# ---- lamch: zml's values are IEEE facts (lamch.zig:15-78 over std.math)
@pytest.mark.parametrize("dt,fn", [(np.float64, "lamch"), (np.float32, "slamch")])
def test_lamch_is_pinned_by_ieee_constants(oracle_lib, dt, fn):
    fi = np.finfo(dt)
    f = getattr(O, fn)
    assert f("E") == fi.eps / 2            # rnd = 1: eps = floatEps / 2
    assert f("P") == fi.eps
    assert f("S") == fi.tiny                # 1/max < tiny, so sfmin = floatMin
    assert f("S") == (2.0 ** -1022 if dt == np.float64 else 2.0 ** -126)
    assert 1.0 / fi.max < fi.tiny
    assert f("U") == fi.tiny
    assert f("O") == fi.max
    assert f("B") == 2 and f("R") == 1
    assert f("N") == fi.nmant               # floatMantissaBits: stored bits (52 / 23)
    assert f("M") == fi.minexp              # floatExponentMin (-1022 / -126)
    assert f("L") == fi.maxexp - 1          # floatExponentMax (1023 / 127)
    assert f("e") == f("E") and f("x") == 0


# ---- the |p| < sfmin branch of getrf2 (getrf2.zig:77-87): divide, not reciprocal
@pytest.mark.parametrize("order", [O.COL, O.ROW])
def test_subnormal_pivot_divides(oracle_lib, order):
    """With p = 2^-1050 the reciprocal overflows (1/p = 2^1050 > max), so only
    the division branch yields the finite multiplier 2^-1060 / p = 2^-10."""
    p, q = 2.0 ** -1050, 2.0 ** -1060
    A = np.array([[p, 1.0, 2.0], [q, 5.0, 0.5], [0.0, 3.0, 1.0]])
    assert p < O.lamch("S") and np.isinf(1.0 / p)
    for fn in (O.getrf, O.getrf2, O.getrf_unblocked):
        a = (A.T if order == O.COL else A).copy().ravel()
        ip = np.zeros(3, np.int32)
        assert fn(order, 3, 3, a, 3, ip) == 0
        LU = a.reshape(3, 3).T if order == O.COL else a.reshape(3, 3)
        assert list(ip) == [1, 2, 3]
        assert LU[1, 0] == 2.0 ** -10          # exact: q / p
        assert np.all(np.isfinite(LU))
    # a planted mistake: the reciprocal branch at this pivot gives an infinite L
    assert np.isinf(q * (1.0 / p))


def test_getrf_info_and_subnormal_singular_column(oracle_lib):
    """A column of subnormals is not singular (p != 0); an exact zero column is."""
    n = 4
    A = np.eye(n)
    A[:, 1] = [0.0, 2.0 ** -1070, 2.0 ** -1072, 0.0]
    a = A.copy().ravel()
    ip = np.zeros(n, np.int32)
    assert O.getrf(O.ROW, n, n, a, n, ip) == 0
    B = np.eye(n)
    B[:, 2] = 0.0
    b = B.copy().ravel()
    assert O.getrf(O.ROW, n, n, b, n, np.zeros(n, np.int32)) == 3


# ---- lange (no zml counterpart): numpy's norms
@pytest.mark.parametrize("order", [O.COL, O.ROW])
@pytest.mark.parametrize("m,n", [(1, 1), (7, 3), (40, 65)])
def test_lange_matches_numpy(oracle_lib, order, m, n):
    A = np.random.default_rng(m * n).uniform(-3, 3, (m, n))
    a = (A.T if order == O.COL else A).copy().ravel()
    ld = m if order == O.COL else n
    assert O.lange(order, "M", m, n, a, ld) == np.abs(A).max()
    assert O.lange(order, "I", m, n, a, ld) == pytest.approx(np.linalg.norm(A, np.inf), rel=1e-15)
    assert O.lange(order, "O", m, n, a, ld) == pytest.approx(np.linalg.norm(A, 1), rel=1e-15)
    assert O.lange(order, "1", m, n, a, ld) == O.lange(order, "O", m, n, a, ld)
    assert O.lange(order, "F", m, n, a, ld) == pytest.approx(np.linalg.norm(A, "fro"), rel=1e-14)
    a[0] = np.nan
    assert np.isnan(O.lange(order, "M", m, n, a, ld)) and np.isnan(O.lange(order, "I", m, n, a, ld))
    assert O.lange(order, "M", 0, n, a, ld) == 0.0


# ---- f32 instantiation: the same generic source at T = f32
class OracleBackendS:
    gemm = staticmethod(O.sgemm)
    trsm = staticmethod(O.strsm)
    syrk = staticmethod(O.ssyrk)


@pytest.mark.parametrize("name", ["gemm", "syrk"])
def test_f32_reference_vectors_exact(oracle_lib, name):
    """The integer-valued gemm/syrk vectors are exact in f32 as well."""
    fails = fixture_replay.replay(name, OracleBackendS, dtype=np.float32)
    assert not fails, "\n".join(fails[:10])


def test_f32_trsm_reference_vectors(oracle_lib):
    """The chained f64 vectors hold at f32 precision: rel 1e-4 (the worst chain
    amplifies the f32 rounding to ~4e-5)."""
    fails = fixture_replay.replay("trsm", OracleBackendS, dtype=np.float32, rel_floor=1e-4)
    assert not fails, "\n".join(fails[:10])


@pytest.mark.parametrize("order", [O.COL, O.ROW])
def test_f32_getrf_closed_form_bitwise(oracle_lib, order):
    n = 150
    a = O.fill_uniform(order, n, n, 11).astype(np.float32)
    b = a.copy()
    ipa, ipb = np.zeros(n, np.int32), np.zeros(n, np.int32)
    assert O.sgetrf(order, n, n, a, n, ipa) == O.sgetrf_unblocked(order, n, n, b, n, ipb) == 0
    assert np.array_equal(ipa, ipb) and np.array_equal(a, b)


def test_f32_potrf_closed_form_bitwise(oracle_lib):
    n = 130
    B = np.random.default_rng(3).standard_normal((n, n))
    S = (B @ B.T / n + np.eye(n)).astype(np.float32)
    a = S.T.copy().ravel()
    b = a.copy()
    assert O.spotrf(O.COL, O.LOWER, n, a, n) == O.spotrf_lower_unblocked(n, b, n) == 0
    L, L2 = np.tril(a.reshape(n, n).T), np.tril(b.reshape(n, n).T)
    assert np.array_equal(L, L2)
    assert np.abs(L.astype(np.float64) @ L.T.astype(np.float64) - S).max() < 1e-5


def test_f32_lu_of_integer_matrix_is_exact(oracle_lib):
    """L = [[1,0,0],[2,1,0],[3,4,1]], U = [[2,1,1],[0,3,1],[0,0,4]]: every
    intermediate of A = L U is an integer, so the f32 factors are exact."""
    L = np.array([[1, 0, 0], [0.5, 1, 0], [0.25, 0.5, 1]])
    U = np.array([[8, 1, 1], [0, 6, 1], [0, 0, 4]])
    A = (L @ U).astype(np.float32)
    a = A.copy().ravel()
    ip = np.zeros(3, np.int32)
    assert O.sgetrf(O.ROW, 3, 3, a, 3, ip) == 0
    assert list(ip) == [1, 2, 3]
    F = a.reshape(3, 3)
    assert np.array_equal(np.tril(F, -1) + np.eye(3), L) and np.array_equal(np.triu(F), U)
