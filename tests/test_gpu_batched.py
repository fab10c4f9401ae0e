"""GPU tests of the batched small solves (zmlhip_batch_plan_*): many
independent systems of mixed sizes in one hipGraph launch, FP32 factor and
FP64 refinement per system.

Systems of the batched configuration: n log-uniform in [8, 256], half general
(U[-1, 1] + n I / 4), half SPD (B B^T / n + I), 1-4 right-hand sides. Every
solution is compared with the oracle's FP64 solve (zml's gesv / posv
restated) and must meet the FP64 backward-error bar; the graph is replayed
several times with eager work in between and every replay is checked.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from tests.helpers import COL, dense, store

pytestmark = pytest.mark.gpu


def make_batch(count, seed, nmin=8, nmax=256, sizes=None):
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    systems = []
    for i, n in enumerate(sizes):
        n = int(n)
        spd = i % 2 == 1
        if spd:
            Bm = rng.standard_normal((n, n))
            A = Bm @ Bm.T / n + np.eye(n)
            A = 0.5 * (A + A.T)
        else:
            A = rng.uniform(-1, 1, (n, n)) + n / 4.0 * np.eye(n)
        nrhs = int(rng.integers(1, 5))
        Bv = rng.uniform(-1, 1, (n, nrhs))
        systems.append((A, Bv, spd))
    return systems


def oracle_solution(A, Bv, spd):
    n, nrhs = Bv.shape
    a = store(COL, A, n)
    b = store(COL, Bv, n)
    if spd:
        assert O.posv(COL, 122, n, nrhs, a, n, b, n) == 0
    else:
        ip = np.zeros(n, np.int32)
        assert O.gesv(COL, n, nrhs, a, n, ip, b, n) == 0
    return dense(COL, n, nrhs, b, n)


class Batch:
    """Device buffers of a batch (one allocation per kind) and its plan."""

    def __init__(self, zh, torch, systems, autotune=1, **opts):
        from zml_amd._lib import IROpts
        self.zh, self.torch, self.systems = zh, torch, systems
        dev = torch.device("cuda")
        na = sum(A.size for A, _, _ in systems)
        nb = sum(Bv.size for _, Bv, _ in systems)
        self.A = torch.empty(max(na, 1), dtype=torch.float64, device=dev)
        self.B = torch.empty(max(nb, 1), dtype=torch.float64, device=dev)
        self.X = torch.zeros(max(nb, 1), dtype=torch.float64, device=dev)
        ha, hb = np.zeros(max(na, 1)), np.zeros(max(nb, 1))
        pa, pb, px, self.offs = [], [], [], []
        oa = ob = 0
        for A, Bv, _ in systems:
            n, nrhs = Bv.shape
            ha[oa:oa + n * n] = store(COL, A, n)
            hb[ob:ob + n * nrhs] = store(COL, Bv, n)
            pa.append(self.A.data_ptr() + 8 * oa)
            pb.append(self.B.data_ptr() + 8 * ob)
            px.append(self.X.data_ptr() + 8 * ob)
            self.offs.append(ob)
            oa += n * n
            ob += n * nrhs
        self.A.copy_(torch.from_numpy(ha))
        self.B.copy_(torch.from_numpy(hb))
        count = len(systems)
        I = ctypes.c_int * max(count, 1)
        Pp = ctypes.c_void_p * max(count, 1)
        self._n = I(*[Bv.shape[0] for _, Bv, _ in systems])
        self._r = I(*[Bv.shape[1] for _, Bv, _ in systems])
        self._s = I(*[1 if spd else 0 for _, _, spd in systems])
        self._a, self._b, self._x = Pp(*pa), Pp(*pb), Pp(*px)
        self.plan = ctypes.c_void_p()
        self.stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        o = IROpts(**opts)
        rc = zh.zmlhip_batch_plan_create(ctypes.byref(self.plan), count, self._n, self._r, self._s, self._a,
                                         self._b, self._x, ctypes.byref(o), autotune, self.stream)
        assert rc == 0, (rc, zh.zmlhip_last_error())

    def run(self):
        assert self.zh.zmlhip_batch_plan_execute(self.plan, self.stream) == 0, self.zh.zmlhip_last_error()
        st = (ctypes.c_int * max(len(self.systems), 1))()
        assert self.zh.zmlhip_batch_plan_status(self.plan, st, self.stream) == 0
        self.torch.cuda.synchronize()
        return np.array(st[:len(self.systems)]), self.X.cpu().numpy()

    def solutions(self, hx):
        for (A, Bv, _), ob in zip(self.systems, self.offs):
            n, nrhs = Bv.shape
            yield dense(COL, n, nrhs, hx[ob:ob + n * nrhs], n)

    def close(self):
        self.zh.zmlhip_batch_plan_destroy(self.plan)


def check(batch, status, hx):
    assert (status >= 0).all(), np.nonzero(status < 0)
    worst_fe = worst_be = 0.0
    for (A, Bv, spd), X in zip(batch.systems, batch.solutions(hx)):
        n = A.shape[0]
        Xo = oracle_solution(A, Bv, spd)
        worst_fe = max(worst_fe, np.abs(X - Xo).max() / np.abs(Xo).max())
        an = np.abs(A).sum(1).max()
        for q in range(Bv.shape[1]):
            be = np.abs(Bv[:, q] - A @ X[:, q]).max() / (an * np.abs(X[:, q]).max() * n * 2.0 ** -53)
            worst_be = max(worst_be, be)
    assert worst_fe < 1e-12 and worst_be <= 16, (worst_fe, worst_be)
    return worst_fe, worst_be


def test_batched_config5_kind(zh, torch_cuda):
    systems = make_batch(300, 47)
    bt = Batch(zh, torch_cuda, systems)
    tg, r, cnt = (ctypes.c_int * 4)(), (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
    assert zh.zmlhip_batch_plan_shapes(bt.plan, tg, r, cnt) == 0
    assert sum(cnt) == 300
    for b in range(4):
        assert tg[b] * r[b] >= (32, 64, 128, 256)[b]
    status, hx = bt.run()
    fe, be = check(bt, status, hx)
    print(f"buckets {list(cnt)} shapes {[(tg[b], r[b]) for b in range(4)]}: refinement steps max "
          f"{status.max()}, error vs oracle {fe:.3g}, backward error {be:.3g}")
    # replays of the same graph, with eager work on the stream in between
    for rep in range(3):
        bt.X.zero_()
        torch_cuda.cuda.synchronize()
        status2, hx2 = bt.run()
        check(bt, status2, hx2)
        assert np.array_equal(status2, status)
    bt.close()


def test_batched_edge_sizes(zh, torch_cuda):
    sizes = [1, 2, 8, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256]
    systems = make_batch(len(sizes), 3, sizes=sizes)
    bt = Batch(zh, torch_cuda, systems, autotune=0)
    status, hx = bt.run()
    check(bt, status, hx)
    bt.close()


def test_batched_needs_pivoting(zh, torch_cuda):
    """General systems without the diagonal shift: every column pivots."""
    rng = np.random.default_rng(9)
    systems = []
    for n in (17, 40, 100, 200):
        A = rng.uniform(-1, 1, (n, n))
        systems.append((A, rng.uniform(-1, 1, (n, 2)), False))
    bt = Batch(zh, torch_cuda, systems)
    status, hx = bt.run()
    assert (status >= 0).all()
    for (A, Bv, _), X in zip(systems, bt.solutions(hx)):
        n = A.shape[0]
        an = np.abs(A).sum(1).max()
        for q in range(2):
            assert np.abs(Bv[:, q] - A @ X[:, q]).max() / (an * np.abs(X[:, q]).max() * n * 2.0 ** -53) <= 16
    bt.close()


def test_batched_empty_and_bad_arguments(zh, torch_cuda):
    plan = ctypes.c_void_p()
    s = ctypes.c_void_p(torch_cuda.cuda.current_stream().cuda_stream)
    assert zh.zmlhip_batch_plan_create(ctypes.byref(plan), 0, None, None, None, None, None, None, None, 1, s) == 0
    assert zh.zmlhip_batch_plan_execute(plan, s) == 0
    assert zh.zmlhip_batch_plan_destroy(plan) == 0
    n = (ctypes.c_int * 1)(300)
    r = (ctypes.c_int * 1)(1)
    p = (ctypes.c_void_p * 1)(1)
    assert zh.zmlhip_batch_plan_create(ctypes.byref(plan), 1, n, r, None, p, p, p, None, 1, s) == -3
    assert zh.zmlhip_batch_plan_create(None, 0, None, None, None, None, None, None, None, 1, s) == -1
