"""GPU tests of the mixed-precision LU solve with FP64 refinement
(zmlhip_dgesv_mixed), through the C ABI on device buffers.

The bar is the one the FP64 path is held to: normwise backward error
||b - Ax||_inf / (||A||_inf ||x||_inf n u) <= 16 (u = 2^-53), and, where the
problem is well conditioned, agreement with the oracle's FP64 solution
(zml's getrf/getrs restated) to 1e-10 relative. Every case reports the error
against the oracle and the refinement counts.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, ROW, P, dense, store

pytestmark = pytest.mark.gpu


def config3_matrix(n, delta, seed, scaled=True):
    """A = diag(r) (R + delta n I) diag(c), R ~ U[-0.5, 0.5], r, c = 10^U(-3, 3)."""
    R = O.fill_uniform(ROW, n, n, seed, -0.5, 0.5).reshape(n, n)
    A = R + delta * n * np.eye(n)
    if scaled:
        r = 10.0 ** O.fill_uniform(ROW, n, 1, seed + 3, -3.0, 3.0)
        c = 10.0 ** O.fill_uniform(ROW, n, 1, seed + 4, -3.0, 3.0)
        A = r[:, None] * A * c[None, :]
    return A


def oracle_solve(A, B):
    n, nrhs = B.shape
    a = A.ravel().copy()
    ip = np.zeros(n, np.int32)
    b = B.copy().ravel()
    assert O.gesv(ROW, n, nrhs, a, n, ip, b, nrhs) == 0
    return b.reshape(n, nrhs)


def run_mixed(zh, torch, A, B, order, precision, **opts):
    from zml_amd._lib import IROpts, IRResult
    n, nrhs = B.shape
    dev = torch.device("cuda")
    a = torch.from_numpy(store(order, A, n)).to(dev)
    ldb = n if order == COL else nrhs
    b = torch.from_numpy(store(order, B, ldb)).to(dev)
    x = torch.zeros_like(b)
    o = IROpts(precision=precision, **opts)
    r = IRResult()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = zh.zmlhip_dgesv_mixed(order, n, nrhs, P(a), n, P(b), ldb, P(x), ldb, ctypes.byref(o),
                               ctypes.byref(r), s)
    assert rc == 0, (rc, zh.zmlhip_last_error())
    torch.cuda.synchronize()
    # A must be unchanged
    assert np.array_equal(a.cpu().numpy(), store(order, A, n))
    return dense(order, n, nrhs, x.cpu().numpy(), ldb), r


def backward_error(A, X, B):
    n = A.shape[0]
    be = []
    for j in range(B.shape[1]):
        r = B[:, j] - A @ X[:, j]
        be.append(np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(X[:, j]).max() * n * 2.0 ** -53))
    return max(be)


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("delta", [1.0, 0.1, 0.01])
def test_mixed_lu_config3_backward_error(zh, torch_cuda, precision, order, delta):
    n = 1000
    A = config3_matrix(n, delta, 100)
    X0 = O.fill_uniform(ROW, n, 1, 43).reshape(n, 1)
    B = A @ X0
    X, res = run_mixed(zh, torch_cuda, A, B, order, precision)
    be = backward_error(A, X, B)
    Xo = oracle_solve(A, B)
    fe = np.abs(X - Xo).max() / np.abs(Xo).max()
    print(f"prec={precision} order={order} delta={delta}: backward error {be:.3g} (n u units), "
          f"lib {res.backward_error:.3g}, IR {res.iterations}, GMRES {res.gmres_iterations}, "
          f"fallback {res.fallback}, error vs oracle {fe:.3g}, "
          f"diagonal-block panels {res.panels_optimistic}/{res.panels}")
    assert res.info == 0
    assert be <= 16 and res.backward_error <= 16
    assert res.fallback < 2
    if precision == 0:   # an FP32 factor refines in a few steps without GMRES
        assert res.iterations <= 3 and res.gmres_iterations == 0
    # these matrices are diagonally heavy after equilibration: no pivot comes from below a
    # block (n = 1000: one 512-wide block column is tried, the 488-wide tail is not)
    assert res.panels == 1 and res.panels_optimistic == 1


@pytest.mark.parametrize("precision,order,n", [(1, ROW, 2560), (2, COL, 2600), (1, COL, 2100)])
def test_mixed_lu_several_block_columns_ahead(zh, torch_cuda, precision, order, n):
    """Orders with four to six 512-wide block columns: every block column's
    update of the next one, handed to the panel stream, and of the rest, with
    the panels' interchanges composed on their second stream. Checked by the
    FP64 backward error and the known solution."""
    A = config3_matrix(n, 0.1, 11)
    X0 = O.fill_uniform(ROW, n, 1, 45).reshape(n, 1)
    B = A @ X0
    X, res = run_mixed(zh, torch_cuda, A, B, order, precision)
    be = backward_error(A, X, B)
    cond_fe = np.abs(X - X0).max() / np.abs(X0).max()
    print(f"n={n} prec={precision}: backward error {be:.3g}, IR {res.iterations}, GMRES {res.gmres_iterations}, "
          f"fallback {res.fallback}, panels {res.panels_optimistic}/{res.panels}, error vs x0 {cond_fe:.3g}")
    assert res.info == 0 and res.fallback == 0
    assert be <= 16 and res.backward_error <= 16
    assert res.iterations <= 5 and res.gmres_iterations == 0
    assert cond_fe < 1e-8


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_mixed_lu_well_conditioned_matches_oracle(zh, torch_cuda, precision):
    n, nrhs = 777, 3   # not a multiple of the block size
    A = config3_matrix(n, 0.1, 7, scaled=False)
    Xt = np.random.default_rng(2).uniform(-1, 1, (n, nrhs))
    B = A @ Xt
    X, res = run_mixed(zh, torch_cuda, A, B, ROW, precision)
    Xo = oracle_solve(A, B)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-10
    assert backward_error(A, X, B) <= 16


def test_mixed_lu_random_uniform_needs_pivoting(zh, torch_cuda):
    """Config 1's kind of matrix: partial pivoting is exercised in every panel."""
    n = 1024
    A = dense(ROW, n, n, O.fill_uniform(ROW, n, n, 42), n)
    X0 = O.fill_uniform(ROW, n, 1, 43).reshape(n, 1)
    B = A @ X0
    for prec in (0, 1, 2):
        X, res = run_mixed(zh, torch_cuda, A, B, COL, prec)
        assert backward_error(A, X, B) <= 16, res.backward_error
        Xo = oracle_solve(A, B)
        assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-9
        assert res.panels == 2 and res.panels_optimistic < 2
    # at threshold 1 (only partial pivoting's own choices) every diagonal-block try
    # with rows below a 256-wide half is rejected and redone
    X, res = run_mixed(zh, torch_cuda, A, B, COL, 1, pivot_threshold=1.0)
    assert res.panels == 2 and res.panels_optimistic == 0
    assert backward_error(A, X, B) <= 16


@pytest.mark.parametrize("mode", ["threshold16", "exact", "full_search"])
def test_mixed_lu_panel_paths_agree(zh, torch_cuda, mode):
    """The same system through diagonal-block panels (threshold 16: all accepted;
    threshold 1, i.e. only where partial pivoting would agree: some rejected and
    redone) and through full pivot searches only."""
    n = 1280
    A = config3_matrix(n, 0.1, 21)
    X0 = O.fill_uniform(ROW, n, 1, 44).reshape(n, 1)
    B = A @ X0
    kw = {"threshold16": {}, "exact": {"pivot_threshold": 1.0}, "full_search": {"no_optimistic": 1}}[mode]
    X, res = run_mixed(zh, torch_cuda, A, B, ROW, 1, **kw)
    print(f"{mode}: diagonal-block panels {res.panels_optimistic}/{res.panels}, IR {res.iterations}, "
          f"GMRES {res.gmres_iterations}")
    # n = 1280: block columns of 512, 512 and 256 (the last one has no rows below it)
    assert res.panels == (0 if mode == "full_search" else 3)
    if mode == "threshold16":
        assert res.panels_optimistic == 3
    if mode == "exact":
        assert 1 <= res.panels_optimistic < 3
    assert backward_error(A, X, B) <= 16


@pytest.mark.parametrize("order", [ROW, COL])
def test_mixed_gmres_and_fp64_fallbacks(zh, torch_cuda, order):
    n = 600
    A = config3_matrix(n, 0.01, 11)
    B = A @ np.ones((n, 1))
    # one classic refinement step only (BF16 needs more): GMRES-IR must finish the job
    X, res = run_mixed(zh, torch_cuda, A, B, order, 2, max_iter=1)
    assert res.iterations <= 1
    assert res.fallback == 1 and res.gmres_iterations > 0
    assert backward_error(A, X, B) <= 16
    # neither: the FP64 factorization solves it
    X, res = run_mixed(zh, torch_cuda, A, B, order, 2, max_iter=1, no_gmres=1)
    assert res.iterations <= 1
    assert res.fallback == 2
    assert backward_error(A, X, B) <= 16


@pytest.mark.parametrize("order", [ROW, COL])
def test_mixed_lu_singular_matrix_is_reported(zh, torch_cuda, order):
    """An exactly singular A (zero column 100): the low-precision factor and
    the FP64 fallback both find the zero pivot; the call returns its column
    (LAPACK's info > 0) and the backward error is not a small number."""
    from zml_amd._lib import IROpts, IRResult
    torch = torch_cuda
    n = 600
    A = config3_matrix(n, 0.1, 31, scaled=False)
    A[:, 100] = 0.0
    B = A @ np.ones((n, 1))
    a = torch.from_numpy(store(order, A, n)).cuda()
    ldb = n if order == COL else 1
    b = torch.from_numpy(store(order, B, ldb)).cuda()
    x = torch.zeros_like(b)
    r = IRResult()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = zh.zmlhip_dgesv_mixed(order, n, 1, P(a), n, P(b), ldb, P(x), ldb,
                               ctypes.byref(IROpts(precision=1, max_iter=3)), ctypes.byref(r), s)
    torch.cuda.synchronize()
    print(f"rc {rc}, info {r.info}, fallback {r.fallback}, be {r.backward_error}")
    assert r.info == 101 and r.fallback == 2
    # a column-major A reaches the FP64 factorization as A^T (rows contiguous),
    # whose zero pivot (the zero row) shows at the last step
    assert rc == (101 if order == ROW else n)
    assert not (r.backward_error <= 16)   # NaN or large, never a small number


def test_mixed_rejects_bad_arguments(zh, torch_cuda):
    from zml_amd._lib import IRResult
    r = IRResult()
    s = ctypes.c_void_p(torch_cuda.cuda.current_stream().cuda_stream)
    assert zh.zmlhip_dgesv_mixed(7, 4, 1, None, 4, None, 4, None, 4, None, ctypes.byref(r), s) == -1
    assert zh.zmlhip_dgesv_mixed(COL, 4, 1, None, 3, None, 4, None, 4, None, ctypes.byref(r), s) == -5
