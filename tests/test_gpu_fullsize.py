"""The configurations at their full sizes on one MI355X, through the C ABI.

Matrices are generated on the device (the oracle cannot factor n = 32768 in
test time); the checks are size-independent properties computed in FP64 on
the device, independently of the library: the normwise backward error
||b - A x||_inf / (||A||_inf ||x||_inf n u) <= 16 (u = 2^-53) of every right-
hand side, the error against the known x_true where the system is well
conditioned, and the permutation validity of the pivots. Every case prints its
time, refinement counts and fallback.

* mixed LU + IR, n = 32768: A = diag(r)(R + delta n I)diag(c), delta in
  {1, 0.1, 0.01}, r, c = 10^U(-3, 3), FP16 / BF16 / FP32 trailing updates;
* mixed Cholesky + IR, n = 32768: A = B B^T / n + mu I, mu in {1, 1e-2, 1e-4},
  16 right-hand sides;
* plain FP64 dgetrf + dgetrs and dpotrf + dpotrs at n = 16384;
* the 20,000-system batch (n log-uniform in [8, 256], half SPD, 1-4 RHS),
  every system's backward error checked.
"""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def config3_base(torch_cuda):
    """R (n x n, U[-0.5, 0.5]) and the scales, shared by the delta cases."""
    torch = torch_cuda
    n = 32768
    g = torch.Generator(device="cuda")
    g.manual_seed(45)
    R = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g).sub_(0.5)
    r = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 6 - 3)
    c = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 6 - 3)
    x_true = torch.rand(n, 1, dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)
    yield R, r, c, x_true
    del R
    torch.cuda.empty_cache()


def _backward_error(torch, A, x, b, anorm):
    res = b - A @ x
    n = A.shape[0]
    return (res.abs().amax(0) / (anorm * x.abs().amax(0) * n * U)).max().item()


@pytest.mark.parametrize("delta", [1.0, 0.1, 0.01])
@pytest.mark.parametrize("precision", [1, 2, 0])
def test_mixed_lu_config3_n32768(zh, torch_cuda, config3_base, delta, precision):
    from zml_amd._lib import IROpts, IRResult
    torch = torch_cuda
    R, r, c, x_true = config3_base
    n = R.shape[0]
    A = R.clone()
    A.diagonal().add_(delta * n)
    A.mul_(r[:, None]).mul_(c[None, :])
    b = A @ x_true
    x = torch.zeros_like(b)
    anorm = A.abs().sum(1).max().item()
    res = IRResult()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = zh.zmlhip_dgesv_mixed(101, n, 1, _p(A), n, _p(b), 1, _p(x), 1, ctypes.byref(IROpts(precision=precision)),
                               ctypes.byref(res), _stream(torch))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert rc == 0, zh.zmlhip_last_error()
    be = _backward_error(torch, A, x, b, anorm)
    err = ((x - x_true).abs().max() / x_true.abs().max()).item()
    print(f"delta {delta} prec {precision}: {dt * 1e3:.1f} ms, IR {res.iterations}, GMRES {res.gmres_iterations}, "
          f"fallback {res.fallback}, diagonal-block panels {res.panels_optimistic}/{res.panels}, "
          f"backward error {be:.3g} (library {res.backward_error:.3g}), error vs x_true {err:.2e}")
    assert be <= 16 and res.backward_error <= 16
    del A, b, x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("mu", [1.0, 1e-2, 1e-4])
def test_mixed_cholesky_config4_n32768(zh, torch_cuda, mu):
    from zml_amd._lib import IROpts, IRResult
    torch = torch_cuda
    n, nrhs = 32768, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(44)
    Bm = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = Bm @ Bm.T
    del Bm
    A.div_(n)
    A.diagonal().add_(mu)
    A = 0.5 * (A + A.T)
    Bt = torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)   # column-major n x 16
    Xt = torch.zeros_like(Bt)
    anorm = A.abs().sum(1).max().item()
    res = IRResult()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = zh.zmlhip_dposv_mixed(102, b"L", n, nrhs, _p(A), n, _p(Bt), n, _p(Xt), n, ctypes.byref(IROpts(precision=1)),
                               ctypes.byref(res), _stream(torch))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert rc == 0, zh.zmlhip_last_error()
    R = Bt - Xt @ A
    be = (R.abs().amax(1) / (anorm * Xt.abs().amax(1) * n * U)).max().item()
    print(f"mu {mu}: {dt * 1e3:.1f} ms, IR {res.iterations}, GMRES {res.gmres_iterations}, fallback "
          f"{res.fallback}, low-precision info {res.info}, worst backward error {be:.3g}")
    assert be <= 16 and res.backward_error <= 16
    del A, Bt, Xt, R
    torch.cuda.empty_cache()


def test_plain_fp64_lu_n16384(zh, torch_cuda):
    torch = torch_cuda
    n = 16384
    g = torch.Generator(device="cuda")
    g.manual_seed(42)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)
    x_true = torch.rand(n, 1, dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)
    b = A @ x_true
    LU = A.clone()
    x = b.clone()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert zh.zmlhip_dgetrf(101, n, n, _p(LU), n, _p(ipiv), _p(info), _stream(torch)) == 0
    assert zh.zmlhip_dgetrs(101, b"N", n, 1, _p(LU), n, _p(ipiv), _p(x), 1, _stream(torch)) == 0
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert info.item() == 0
    ip = ipiv.cpu().numpy()
    assert ((ip >= np.arange(1, n + 1)) & (ip <= n)).all()   # ipiv[i] >= i + 1, <= n
    be = _backward_error(torch, A, x, b, A.abs().sum(1).max().item())
    err = ((x - x_true).abs().max() / x_true.abs().max()).item()
    print(f"dgetrf+dgetrs n={n}: {dt * 1e3:.1f} ms ({2 / 3 * n ** 3 / dt / 1e12:.1f} TF/s), backward error "
          f"{be:.3g}, error vs x_true {err:.2e}")
    # U[-1, 1] at n = 16384 has a condition number of about 1e5: a backward
    # error of ~1e-2 (in units of n u) then allows a forward error of ~1e-8
    assert be <= 16 and err < 1e-7


def test_plain_fp64_cholesky_n16384(zh, torch_cuda):
    torch = torch_cuda
    n, nrhs = 16384, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(44)
    Bm = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = Bm @ Bm.T
    del Bm
    A.div_(n)
    A.diagonal().add_(1.0)
    A = 0.5 * (A + A.T)
    X_true = torch.rand(nrhs, n, dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)
    Bt = X_true @ A
    L = A.clone()
    Xt = Bt.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert zh.zmlhip_dpotrf(102, b"L", n, _p(L), n, _p(info), _stream(torch)) == 0
    assert zh.zmlhip_dpotrs(102, b"L", n, nrhs, _p(L), n, _p(Xt), n, _stream(torch)) == 0
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert info.item() == 0
    anorm = A.abs().sum(1).max().item()
    R = Bt - Xt @ A
    be = (R.abs().amax(1) / (anorm * Xt.abs().amax(1) * n * U)).max().item()
    err = ((Xt - X_true).abs().max() / X_true.abs().max()).item()
    print(f"dpotrf+dpotrs n={n}: {dt * 1e3:.1f} ms ({n ** 3 / 3 / dt / 1e12:.1f} TF/s), backward error {be:.3g}, "
          f"error vs x_true {err:.2e}")
    assert be <= 16 and err < 1e-9


def test_batched_20000_systems(zh, torch_cuda):
    """The batched configuration at full size: every system's every right-hand
    side meets the FP64 backward-error bar."""
    from tests.test_gpu_batched import Batch, make_batch
    systems = make_batch(20000, 47)
    bt = Batch(zh, torch_cuda, systems)
    try:
        bt.run()
        t0 = time.perf_counter()
        status, hx = bt.run()
        dt = time.perf_counter() - t0
        assert (status >= 0).all(), np.nonzero(status < 0)
        worst = 0.0
        for (A, Bv, _), X in zip(systems, bt.solutions(hx)):
            n = A.shape[0]
            an = np.abs(A).sum(1).max()
            be = (np.abs(Bv - A @ X).max(0) / (an * np.abs(X).max(0) * n * U)).max()
            worst = max(worst, be)
        print(f"20000 systems: {dt * 1e3:.2f} ms per graph launch (incl. status copy), refinement steps <= "
              f"{status.max()}, worst backward error {worst:.3g}")
        assert worst <= 16
    finally:
        bt.close()
