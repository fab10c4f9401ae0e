"""GPU parity of the remaining LAPACK entries through the C ABI:

* LAPACKE_dgetrf2 / dpotrf2 / dlamch against the oracle's getrf2 / potrf2 /
  lamch (zml's recursive panel, recursive Cholesky and machine constants);
* the |p| < sfmin branch of getrf2.zig:77-87 on the GPU (a subnormal pivot,
  whose reciprocal overflows, must divide);
* LAPACKE_sgetrf / spotrf (zml's getrf / potrf at T = f32) against the
  oracle's f32 instantiation;
* the bounded cross-workgroup wait of the LU panel giving up (forced from the
  host with zmlhip_set_panel_spin_limit, every kernel still ending normally):
  reported as ZMLHIP_ESYNC once, cleared by the next factorization;
* two factorizations queued on two streams by one thread do not share scratch
  memory at the same time (both match the oracle).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, ROW, U64, P, dense, store

pytestmark = pytest.mark.gpu
ESYNC = -1002


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (64, 64), (300, 300), (700, 260), (260, 700)])
def test_dgetrf2_matches_oracle_getrf2(zh, lu_panel, order, m, n):
    a = O.fill_uniform(order, m, n, 500 + m + 3 * n)
    ld = m if order == COL else n
    ref = a.copy()
    ip_ref = np.zeros(min(m, n), np.int32)
    info_ref = O.getrf2(order, m, n, ref, ld, ip_ref)
    ip = np.zeros(min(m, n), np.int32)
    assert zh.LAPACKE_dgetrf2(order, m, n, P(a), ld, P(ip)) == info_ref == 0
    assert np.array_equal(ip, ip_ref)
    got, exp = dense(order, m, n, a, ld), dense(order, m, n, ref, ld)
    assert np.abs(got - exp).max() <= 1e-11 * max(1.0, np.abs(exp).max())


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("n", [1, 2, 63, 300, 700])
def test_dpotrf2_matches_oracle_potrf2(zh, order, uplo, n):
    B = O.fill_uniform(ROW, n, n, 70 + n).reshape(n, n)
    A = B @ B.T / n + np.eye(n)
    a = store(order, A, n)
    ref = a.copy()
    code = 122 if uplo == b"L" else 121
    assert zh.LAPACKE_dpotrf2(order, uplo, n, P(a), n) == O.potrf2(order, code, n, ref, n) == 0
    tri = np.tril if uplo == b"L" else np.triu
    got, exp = dense(order, n, n, a, n), dense(order, n, n, ref, n)
    assert np.abs(tri(got) - tri(exp)).max() < 1e-12
    # a non-positive pivot is reported at the same column as zml's potrf2
    A2 = A.copy()
    A2[n // 2, n // 2] = -1.0
    a2, r2 = store(order, A2, n), store(order, A2, n)
    assert zh.LAPACKE_dpotrf2(order, uplo, n, P(a2), n) == O.potrf2(order, code, n, r2, n) == n // 2 + 1


def test_dlamch_matches_oracle_for_every_character(zh):
    for c in "EeSsBbPpNnRrMmUuLlOoXq":
        assert zh.LAPACKE_dlamch(c.encode()) == O.lamch(c), c
    assert zh.LAPACKE_dlamch(b"S") == 2.0 ** -1022 and zh.LAPACKE_dlamch(b"N") == 52


@pytest.mark.parametrize("order", [ROW, COL])
def test_subnormal_pivot_divides_on_the_gpu(zh, lu_panel, order):
    """p = 2^-1050 < sfmin: 1/p overflows, so only division gives l = 2^-10."""
    p, q = 2.0 ** -1050, 2.0 ** -1060
    A = np.array([[p, 1.0, 2.0], [q, 5.0, 0.5], [0.0, 3.0, 1.0]])
    a = store(order, A, 3)
    ip = np.zeros(3, np.int32)
    assert zh.LAPACKE_dgetrf(order, 3, 3, P(a), 3, P(ip)) == 0
    LU = dense(order, 3, 3, a, 3)
    assert list(ip) == [1, 2, 3] and LU[1, 0] == 2.0 ** -10 and np.all(np.isfinite(LU))


@pytest.mark.parametrize("order", [ROW, COL])
def test_subnormal_pivot_column_in_a_multi_workgroup_panel(zh, lu_panel, order):
    """A column of subnormals in the middle of a tall matrix (several panel
    workgroups): the GPU leaf divides exactly where zml's getrf2 does, so the
    factors equal the oracle's there and the pivots are identical."""
    n = 700
    rng = np.random.default_rng(21)
    A = rng.uniform(-1, 1, (n, n))
    k = 300
    # after k eliminations column k is a subnormal multiple of a fixed vector:
    # make it subnormal from the start and keep it out of the earlier updates
    A[:, k] = 0.0
    A[k:, k] = rng.uniform(0.5, 1.0, n - k) * 2.0 ** -1050
    A[k:, :k] = 0.0          # block upper triangular: column k sees no update
    A[np.arange(k), np.arange(k)] += 4.0
    a = store(order, A, n)
    ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    assert O.getrf(order, n, n, ref, n, ip_ref) == 0
    ip = np.zeros(n, np.int32)
    assert zh.LAPACKE_dgetrf(order, n, n, P(a), n, P(ip)) == 0
    assert np.array_equal(ip, ip_ref)
    got, exp = dense(order, n, n, a, n), dense(order, n, n, ref, n)
    # the multipliers of the subnormal column: quotients of subnormals, exact
    assert np.array_equal(got[k + 1:, k], exp[k + 1:, k])
    assert np.all(np.isfinite(got)) and np.abs(got[k + 1:, k]).max() <= 1.0


# ---- FP32 LU / Cholesky -----------------------------------------------------------
@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("n", [3, 64, 200, 600])
def test_sgetrf_matches_f32_oracle(zh, order, n):
    a64 = O.fill_uniform(order, n, n, 900 + n)
    gap = O.min_pivot_gap(order, n, a64, n)
    a = a64.astype(np.float32)
    ref = a.copy()
    ip_ref = np.zeros(n, np.int32)
    assert O.sgetrf(order, n, n, ref, n, ip_ref) == 0
    ip = np.zeros(n, np.int32)
    assert zh.LAPACKE_sgetrf(order, n, n, P(a), n, P(ip)) == 0
    if gap > 1e-3:   # pivots are decided by margins far above f32 rounding
        assert np.array_equal(ip, ip_ref)
    # P A = L U in f32 accuracy, for whatever pivots were taken
    A = dense(order, n, n, a64.astype(np.float32).astype(np.float64), n)
    LU = dense(order, n, n, a.astype(np.float64), n)
    L, U = np.tril(LU, -1) + np.eye(n), np.triu(LU)
    perm = np.arange(n)
    for i, p in enumerate(ip):
        perm[[i, p - 1]] = perm[[p - 1, i]]
    res = np.abs(A[perm] - L @ U).max() / np.abs(A).max()
    assert res <= 4 * n * 2.0 ** -24, res


def test_sgetrf_exact_integer_factors(zh):
    L = np.array([[1, 0, 0], [0.5, 1, 0], [0.25, 0.5, 1]])
    U = np.array([[8, 1, 1], [0, 6, 1], [0, 0, 4]])
    a = (L @ U).astype(np.float32).ravel()
    ip = np.zeros(3, np.int32)
    assert zh.LAPACKE_sgetrf(ROW, 3, 3, P(a), 3, P(ip)) == 0
    F = a.reshape(3, 3).astype(np.float64)
    assert list(ip) == [1, 2, 3]
    assert np.array_equal(np.tril(F, -1) + np.eye(3), L) and np.array_equal(np.triu(F), U)


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("n", [1, 130, 700])
def test_spotrf_matches_f32_oracle(zh, order, uplo, n):
    B = np.random.default_rng(n).standard_normal((n, n))
    A = (B @ B.T / n + np.eye(n)).astype(np.float32)
    a = store(order, A.astype(np.float64), n).astype(np.float32)
    ref = a.copy()
    code = 122 if uplo == b"L" else 121
    assert zh.LAPACKE_spotrf(order, uplo, n, P(a), n) == O.spotrf(order, code, n, ref, n) == 0
    tri = np.tril if uplo == b"L" else np.triu
    got = tri(dense(order, n, n, a.astype(np.float64), n))
    exp = tri(dense(order, n, n, ref.astype(np.float64), n))
    assert np.abs(got - exp).max() <= 64 * n * 2.0 ** -24 * max(1.0, np.abs(exp).max())
    A2 = A.copy()
    A2[n // 2, n // 2] = -1.0
    a2 = store(order, A2.astype(np.float64), n).astype(np.float32)
    assert zh.LAPACKE_spotrf(order, uplo, n, P(a2), n) >= 1


# ---- the panel's bounded wait giving up ---------------------------------------------
def test_panel_give_up_is_reported_once_and_cleared(zh, lu_panel, torch_cuda):
    torch = torch_cuda
    n = 900   # four 256-row panel workgroups: the exchange is used
    a0 = O.fill_uniform(ROW, n, n, 77)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ip = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    try:
        assert zh.zmlhip_set_panel_spin_limit(-1) == 0
        a = torch.from_numpy(a0.copy()).cuda()
        assert zh.zmlhip_dgetrf(ROW, n, n, P(a), n, P(ip), P(info), s) == 0
        torch.cuda.synchronize()
        assert info.item() == ESYNC
        ipv = ip.cpu().numpy()   # the pivots stay inside the matrix
        assert ipv.min() >= 1 and ipv.max() <= n
        a_host = a0.copy()
        iph = np.zeros(n, np.int32)
        assert zh.LAPACKE_dgetrf(ROW, n, n, P(a_host), n, P(iph)) == ESYNC
        assert "gave up" in zh.zmlhip_last_error().decode()
    finally:
        zh.zmlhip_set_panel_spin_limit(0)
    # the next factorization starts clean and is right
    a = torch.from_numpy(a0.copy()).cuda()
    assert zh.zmlhip_dgetrf(ROW, n, n, P(a), n, P(ip), P(info), s) == 0
    torch.cuda.synchronize()
    assert info.item() == 0
    ref = a0.copy()
    ip_ref = np.zeros(n, np.int32)
    O.getrf(ROW, n, n, ref, n, ip_ref)
    assert np.array_equal(ip.cpu().numpy(), ip_ref)
    a_host = a0.copy()
    assert zh.LAPACKE_dgetrf(ROW, n, n, P(a_host), n, P(iph)) == 0


def test_mixed_solver_reports_a_panel_give_up(zh, torch_cuda):
    from zml_amd._lib import IROpts, IRResult
    torch = torch_cuda
    n = 1200
    A = O.fill_uniform(ROW, n, n, 78).reshape(n, n) + 2 * np.eye(n)
    a = torch.from_numpy(A.ravel().copy()).cuda()
    b = torch.from_numpy(A @ np.ones(n)).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    res = IRResult()
    try:
        zh.zmlhip_set_panel_spin_limit(-1)
        rc = zh.zmlhip_dgesv_mixed(ROW, n, 1, P(a), n, P(b), 1, P(x), 1,
                                   ctypes.byref(IROpts(precision=0, no_optimistic=1)), ctypes.byref(res), s)
        assert rc == ESYNC
    finally:
        zh.zmlhip_set_panel_spin_limit(0)
    rc = zh.zmlhip_dgesv_mixed(ROW, n, 1, P(a), n, P(b), 1, P(x), 1,
                               ctypes.byref(IROpts(precision=0, no_optimistic=1)), ctypes.byref(res), s)
    assert rc == 0 and res.backward_error <= 16
    assert np.abs(x.cpu().numpy() - 1).max() < 1e-10


# ---- one thread, two streams ------------------------------------------------------
def test_two_streams_do_not_share_scratch_concurrently(zh, lu_panel, torch_cuda):
    torch = torch_cuda
    n1, n2 = 1500, 1100
    a1h, a2h = O.fill_uniform(ROW, n1, n1, 81), O.fill_uniform(ROW, n2, n2, 82)
    refs = []
    for n, ah in ((n1, a1h), (n2, a2h)):
        r = ah.copy()
        ip = np.zeros(n, np.int32)
        O.getrf(ROW, n, n, r, n, ip)
        refs.append((r, ip))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    a1, a2 = torch.from_numpy(a1h.copy()).cuda(), torch.from_numpy(a2h.copy()).cuda()
    ip1 = torch.zeros(n1, dtype=torch.int32, device="cuda")
    ip2 = torch.zeros(n2, dtype=torch.int32, device="cuda")
    info = torch.zeros(2, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        assert zh.zmlhip_dgetrf(ROW, n1, n1, P(a1), n1, P(ip1), ctypes.c_void_p(info.data_ptr()),
                                ctypes.c_void_p(s1.cuda_stream)) == 0
        assert zh.zmlhip_dgetrf(ROW, n2, n2, P(a2), n2, P(ip2), ctypes.c_void_p(info.data_ptr() + 4),
                                ctypes.c_void_p(s2.cuda_stream)) == 0
        torch.cuda.synchronize()
        assert info.tolist() == [0, 0]
        for (r, ipr), a, ip, n in ((refs[0], a1, ip1, n1), (refs[1], a2, ip2, n2)):
            assert np.array_equal(ip.cpu().numpy(), ipr)
            assert np.abs(a.cpu().numpy() - r).max() <= 1e-10
        a1.copy_(torch.from_numpy(a1h))
        a2.copy_(torch.from_numpy(a2h))
        torch.cuda.synchronize()
