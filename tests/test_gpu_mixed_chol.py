"""GPU tests of the mixed-precision Cholesky solve with FP64 refinement
(zmlhip_dposv_mixed), through the C ABI on device buffers.

Matrices of the mixed-Cholesky configuration: A = B B^T / n + mu I, B with
i.i.d. N(0, 1) entries, mu in {1, 1e-2, 1e-4} (condition about (4 + mu) / mu),
16 right-hand sides uniform in [-1, 1]. Every right-hand side must reach the
FP64 bar ||b - Ax||_inf / (||A||_inf ||x||_inf n u) <= 16, and the solution
is compared with the oracle's FP64 posv (zml's potrf/potrs restated); each
case prints its backward error, refinement counts and error vs the oracle.
Only the `uplo` triangle of A is handed over: the other one holds garbage.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, ROW, P, dense, store

pytestmark = pytest.mark.gpu


def spd_matrix(n, mu, seed):
    rng = np.random.default_rng(seed)
    Bm = rng.standard_normal((n, n))
    A = Bm @ Bm.T / n + mu * np.eye(n)
    return 0.5 * (A + A.T)


def with_garbage(A, uplo):
    """A copy whose unreferenced triangle is overwritten."""
    G = A.copy()
    iu = np.triu_indices(A.shape[0], 1)
    if uplo == b"L":
        G[iu] = 1e30
    else:
        G[(iu[1], iu[0])] = -1e30
    return G


def oracle_posv(A, B):
    """zml's posv, column-major lower (122): row-major would hit zml's potrs
    ldb check (potrs.zig:30 wants ldb >= n in both layouts), which leaves b
    unsolved for ldb = nrhs."""
    n, nrhs = B.shape
    a = store(COL, A, n)
    b = store(COL, B, n)
    assert O.posv(COL, 122, n, nrhs, a, n, b, n) == 0
    return dense(COL, n, nrhs, b, n)


def run_posv_mixed(zh, torch, A, B, order, uplo, precision, **opts):
    from zml_amd._lib import IROpts, IRResult
    n, nrhs = B.shape
    dev = torch.device("cuda")
    a_stored = store(order, with_garbage(A, uplo), n)
    a = torch.from_numpy(a_stored).to(dev)
    ldb = n if order == COL else nrhs
    b = torch.from_numpy(store(order, B, ldb)).to(dev)
    x = torch.zeros_like(b)
    o = IROpts(precision=precision, **opts)
    r = IRResult()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = zh.zmlhip_dposv_mixed(order, uplo, n, nrhs, P(a), n, P(b), ldb, P(x), ldb, ctypes.byref(o),
                               ctypes.byref(r), s)
    assert rc == 0, (rc, zh.zmlhip_last_error())
    torch.cuda.synchronize()
    assert np.array_equal(a.cpu().numpy(), a_stored)   # A untouched
    return dense(order, n, nrhs, x.cpu().numpy(), ldb), r


def backward_errors(A, X, B):
    n = A.shape[0]
    an = np.abs(A).sum(1).max()
    return np.array([np.abs(B[:, j] - A @ X[:, j]).max() / (an * np.abs(X[:, j]).max() * n * 2.0 ** -53)
                     for j in range(B.shape[1])])


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("mu", [1.0, 1e-2, 1e-4])
def test_mixed_cholesky_config4(zh, torch_cuda, precision, mu):
    n, nrhs = 1100, 16        # not a multiple of the 256 block
    A = spd_matrix(n, mu, 5)
    B = np.random.default_rng(44).uniform(-1, 1, (n, nrhs))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, ROW, b"L", precision)
    be = backward_errors(A, X, B)
    Xo = oracle_posv(A, B)
    fe = np.abs(X - Xo).max() / np.abs(Xo).max()
    print(f"prec={precision} mu={mu}: worst backward error {be.max():.3g} (lib {res.backward_error:.3g}), "
          f"IR {res.iterations}, GMRES {res.gmres_iterations}, fallback {res.fallback}, info {res.info}, "
          f"error vs oracle {fe:.3g}")
    assert be.max() <= 16 and res.backward_error <= 16
    assert res.fallback < 2
    assert fe < 1e-13 * (4 + mu) / mu * 1e3   # kappa-scaled agreement with the FP64 oracle
    # the preconditioner must be the Cholesky factor it claims to be: an FP32 factor of
    # these condition numbers (<= 4e4) refines in a few steps without GMRES
    if precision == 0:
        assert res.iterations <= 3 and res.gmres_iterations == 0


@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_mixed_cholesky_layouts(zh, torch_cuda, order, uplo):
    n, nrhs = 700, 5
    A = spd_matrix(n, 1e-2, 9)
    B = np.random.default_rng(45).uniform(-1, 1, (n, nrhs))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, order, uplo, 1)
    assert backward_errors(A, X, B).max() <= 16
    Xo = oracle_posv(A, B)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-9


@pytest.mark.parametrize("n", [333, 130, 1])
def test_mixed_cholesky_unaligned_orders(zh, torch_cuda, n):
    """Orders that are not multiples of 4 (scalar load paths of the solves and
    the residual), a single partial block and the 1 x 1 system."""
    nrhs = 7
    A = spd_matrix(n, 1e-2, 21)
    B = np.random.default_rng(48).uniform(-1, 1, (n, nrhs))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, ROW, b"L", 0)
    assert backward_errors(A, X, B).max() <= 16
    if n > 1:
        # (for n = 1 the rounding of x alone leaves a backward error near the default
        # target of 1, so refinement may hand over to GMRES and FP64)
        assert res.iterations <= 3 and res.gmres_iterations == 0
    Xo = oracle_posv(A, B)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-10


def test_mixed_cholesky_many_rhs_groups(zh, torch_cuda):
    """More than one group of 16 right-hand sides."""
    n, nrhs = 520, 37
    A = spd_matrix(n, 1.0, 3)
    B = np.random.default_rng(46).uniform(-1, 1, (n, nrhs))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, COL, b"L", 2)
    assert backward_errors(A, X, B).max() <= 16
    Xo = oracle_posv(A, B)
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-11


@pytest.mark.parametrize("order", [ROW, COL])
def test_mixed_cholesky_gmres_and_fp64_fallbacks(zh, torch_cuda, order):
    n = 600
    A = spd_matrix(n, 1e-4, 12)
    B = np.random.default_rng(47).uniform(-1, 1, (n, 3))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, order, b"L", 2, max_iter=1)
    assert res.fallback >= 1 and res.gmres_iterations > 0
    assert backward_errors(A, X, B).max() <= 16
    X, res = run_posv_mixed(zh, torch_cuda, A, B, order, b"L", 2, max_iter=1, no_gmres=1)
    assert res.fallback == 2
    assert backward_errors(A, X, B).max() <= 16


def test_mixed_cholesky_indefinite_low_precision_gets_shifted(zh, torch_cuda):
    """A matrix that is SPD but so close to singular that the low-precision
    factor loses definiteness: the factorization is repeated with a shift and
    the refinement still reaches the FP64 bar."""
    n = 512
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, -9, n)
    A = (Q * ev) @ Q.T
    A = 0.5 * (A + A.T)
    B = rng.uniform(-1, 1, (n, 2))
    X, res = run_posv_mixed(zh, torch_cuda, A, B, ROW, b"L", 1)
    print(f"info {res.info}, fallback {res.fallback}, GMRES {res.gmres_iterations}, be {res.backward_error:.3g}")
    assert backward_errors(A, X, B).max() <= 16


@pytest.mark.parametrize("order", [ROW, COL])
def test_mixed_cholesky_not_positive_definite_is_reported(zh, torch_cuda, order):
    """A clearly indefinite A: the shifted low-precision factorizations keep
    failing, and the call returns the first non-positive pivot (LAPACK's info)
    with x untouched, as dposv would."""
    from zml_amd._lib import IROpts, IRResult
    torch = torch_cuda
    n = 300
    A = np.eye(n)
    A[10, 10] = -100.0
    B = np.ones((n, 1))
    a = torch.from_numpy(store(order, A, n)).cuda()
    ldb = n if order == COL else 1
    b = torch.from_numpy(store(order, B, ldb)).cuda()
    x = torch.full_like(b, 7.0)
    r = IRResult()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = zh.zmlhip_dposv_mixed(order, b"L", n, 1, P(a), n, P(b), ldb, P(x), ldb, ctypes.byref(IROpts(precision=1)),
                               ctypes.byref(r), s)
    torch.cuda.synchronize()
    assert rc == r.info == 11
    assert np.all(x.cpu().numpy() == 7.0)


def test_mixed_cholesky_rejects_bad_arguments(zh, torch_cuda):
    from zml_amd._lib import IRResult
    r = IRResult()
    s = ctypes.c_void_p(torch_cuda.cuda.current_stream().cuda_stream)
    assert zh.zmlhip_dposv_mixed(7, b"L", 4, 1, None, 4, None, 4, None, 4, None, ctypes.byref(r), s) == -1
    assert zh.zmlhip_dposv_mixed(ROW, b"X", 4, 1, None, 4, None, 4, None, 4, None, ctypes.byref(r), s) == -2
    assert zh.zmlhip_dposv_mixed(COL, b"L", 4, 1, None, 3, None, 4, None, 4, None, ctypes.byref(r), s) == -6
