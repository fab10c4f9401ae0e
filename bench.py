"""Headline benchmark: LU solve TFLOP/s at FP64 backward error, mixed
precision + iterative refinement versus plain FP64, n = 32768, on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n N] [--delta D] [--precision fp16|bf16|fp32]

Workload (the mixed-precision LU configuration): A = diag(r) (R + delta n I)
diag(c), R ~ U[-0.5, 0.5], r, c = 10^U(-3, 3), one right-hand side b = A x_true,
synthetic and generated on the device; A and b are resident in HBM before
timing starts. One step is one complete solve through the C ABI entry point
zmlhip_dgesv_mixed: equilibration, FP32/low-precision LU, triangular solves
and FP64 refinement until ||b - Ax||_inf <= ||A||_inf ||x||_inf n u. A step
only counts if the result meets the FP64 backward-error bar (<= 16); the
throughput is 2/3 n^3 flops per step.

The plain FP64 path (zmlhip_dgetrf + zmlhip_dgetrs on a copy of A) is timed
on the same system for the speed-up. The roofline object describes the
dominant kernel, the low-precision trailing-update GEMM, from HIP events the
library records around each of its launches on the stream it runs on. The
CPU baseline is zml's own algorithm (the C restatement of getrf/getrs in
oracle/) on a bounded sample, one core.

`--workload batched` measures the batched configuration: 20,000 independent
systems, n log-uniform in [8, 256], half general (U[-1, 1] + n I / 4), half
SPD (B B^T / n + I), 1-4 right-hand sides each, solved by one hipGraph
launch of a zmlhip_batch_plan (FP32 factors, FP64 refinement), reported as
systems/s and TFLOP/s (2/3 n^3 general, 1/3 n^3 SPD) next to the fraction of
HBM bandwidth, with zml's gesv/posv (the oracle's C restatement) on all host
cores as the CPU baseline.

`--workload plain` measures the plain FP64 configuration (default n = 16384):
zmlhip_dgetrf + zmlhip_dgetrs on A ~ U[-1, 1] (2/3 n^3), zmlhip_dpotrf +
zmlhip_dpotrs on A = B B^T / n + I (1/3 n^3), one right-hand side each, and
zmlhip_dgemm n x n x n (2 n^3), each step checked (backward error <= 16),
with the FP64 trailing-update GEMM inside the LU as the roofline kernel.

`--workload cholesky` measures the mixed-precision Cholesky configuration
instead: A = B B^T / n + mu I (B with N(0, 1) entries), 16 right-hand sides,
zmlhip_dposv_mixed against the plain FP64 zmlhip_dpotrf + zmlhip_dpotrs, at
1/3 n^3 flops per solve.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LU solve TFLOP/s at FP64 backward error, mixed+IR vs plain FP64, n=32768"
PEAK = {"fp16": 2500.0, "bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}   # dense TFLOP/s, MI355X
PREC = {"fp32": 0, "fp16": 1, "bf16": 2}
PROF_GEMM_F64, PROF_GEMM_LOWP = 1, 2


def cpu_model():
    """The host CPU's model name (lscpu's "Model name"), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sizes=(1024, 2048, 4096), reps=(1, 1, 2), n_target=32768):
    """zml's getrf + getrs (oracle/zml_oracle.c restatement), one core.

    Times the reference's own small configuration (n = 1024) and n = 2048,
    4096 on uniform[-1, 1] row-major systems, checks every solve's normwise
    backward error (<= 16, u = 2^-53), fits t = a n^3 through the two largest
    sizes and labels the n = 32768 time that fit gives "extrapolated" (it is
    not run: the restatement needs most of an hour there). `value` is the rate
    measured at the largest size."""
    import numpy as np
    import oracle as O
    O.lib()
    times, bes = {}, {}
    for n, r in zip(sizes, reps):
        best = None
        for _ in range(r):
            a = O.fill_uniform(O.ROW, n, n, 42)
            b = O.fill_uniform(O.ROW, n, 1, 43)
            a0, b0 = a.copy(), b.copy()
            ip = np.zeros(n, np.int32)
            t0 = time.perf_counter()
            assert O.getrf(O.ROW, n, n, a, n, ip) == 0
            O.getrs(O.ROW, 111, n, 1, a, n, ip, b, 1)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        x = b.reshape(n)
        res = b0.reshape(n) - a0.reshape(n, n) @ x
        be = float(np.abs(res).max() / (np.abs(a0.reshape(n, n)).sum(1).max() * np.abs(x).max() * n * 2.0 ** -53))
        assert be <= 16, f"CPU baseline n={n}: backward error {be}"
        times[n], bes[n] = best, be
    n1, n2 = sizes[-2], sizes[-1]
    coef = (times[n1] / n1 ** 3 + times[n2] / n2 ** 3) / 2.0
    n = sizes[-1]
    return {"value": round((2.0 / 3.0) * n ** 3 / times[n] / 1e12, 6), "unit": "TFLOP/s", "cores": 1,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"zml getrf+getrs restated in C (oracle/zml_oracle.c, NB=64, no FMA), row-major "
                      f"uniform[-1,1], 1 RHS, one core: value = n={n} (best of {reps[-1]}); every solve's "
                      f"backward error <= 16",
            "seconds": {str(k): round(v, 4) for k, v in times.items()},
            "backward_error": {str(k): round(v, 3) for k, v in bes.items()},
            "config_n1024_s": round(times[sizes[0]], 4),
            "fit": {"model": "t = a n^3", "from_n": [n1, n2], "a_s": coef,
                    f"n{n_target}_s": round(coef * n_target ** 3, 1),
                    f"n{n_target}_tflops": round((2.0 / 3.0) * n_target ** 3 / (coef * n_target ** 3) / 1e12, 6),
                    "label": "extrapolated"}}


HBM_PEAK_GBPS = 8000.0


def roofline_of(kernel, ms, launches, flops, bytes_, peak_tflops, steps, traffic):
    """The binding roof of a kernel from its algorithmic work: HBM when its
    arithmetic intensity (flops per algorithmic byte) is below the ridge
    peak_flops / peak_bandwidth, the matrix cores otherwise. `achieved` is the
    algorithmic rate over the summed HIP-event durations of its launches."""
    if ms <= 0 or launches <= 0:
        return None
    t = ms / 1e3
    ai = flops / bytes_ if bytes_ > 0 else float("inf")
    ridge = peak_tflops * 1e12 / (HBM_PEAK_GBPS * 1e9)
    out = {"kernel": kernel, "launches_per_step": launches // max(1, steps),
           "avg_launch_ms": round(ms / launches, 4),
           "algorithmic_flops_per_launch": round(flops / launches),
           "algorithmic_bytes_per_launch": round(bytes_ / launches),
           "flops_per_byte": round(ai, 1), "ridge_flops_per_byte": round(ridge, 1),
           "traffic": traffic, "mfma_tflops": round(flops / t / 1e12, 2), "mfma_peak": peak_tflops}
    if ai < ridge:
        achieved = bytes_ / t / 1e9
        out.update({"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4)})
    else:
        achieved = flops / t / 1e12
        out.update({"bound": "mfma", "achieved": round(achieved, 2), "peak": peak_tflops, "unit": "TFLOP/s",
                    "frac": round(achieved / peak_tflops, 4)})
    return out


def traffic_from_profiles(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None, help="order (32768; 16384 for --workload plain)")
    ap.add_argument("--delta", type=float, default=0.1)
    ap.add_argument("--precision", default="fp16", choices=sorted(PREC))
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-abi", action="store_true",
                    help="plain workload: skip the end-to-end LAPACKE_dgesv timing through host pointers")
    ap.add_argument("--workload", default="lu", choices=["lu", "cholesky", "batched", "plain"])
    ap.add_argument("--systems", type=int, default=20000)
    ap.add_argument("--mu", type=float, default=1e-2)
    ap.add_argument("--nrhs", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the dense factor-and-solve path is built and timed on a "
              "single MI355X (the multi-GPU layout is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.n is None:
        args.n = 16384 if args.workload == "plain" else 32768
    if args.workload == "plain":
        return bench_plain(args)
    if args.workload == "cholesky":
        return bench_cholesky(args)
    if args.workload == "batched":
        return bench_batched(args)
    import torch
    from zml_amd import _lib
    from zml_amd._lib import IROpts, IRResult
    zh = _lib.load()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    n = args.n

    # ---- synthetic system (device-generated, FP64)
    g = torch.Generator(device=dev)
    g.manual_seed(45)
    A = torch.rand(n, n, dtype=torch.float64, device=dev, generator=g).sub_(0.5)
    A.diagonal().add_(args.delta * n)
    r = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device=dev, generator=g) * 6 - 3)
    c = torch.pow(10.0, torch.rand(n, dtype=torch.float64, device=dev, generator=g) * 6 - 3)
    A.mul_(r[:, None]).mul_(c[None, :])
    del r, c
    x_true = torch.rand(n, 1, dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    b = A @ x_true
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    anorm = A.abs().sum(1).max().item()

    def backward_error(xv):
        res = b - A @ xv
        return (res.abs().max() / (anorm * xv.abs().max() * n * 2.0 ** -53)).item()

    # ---- mixed precision + IR (row-major view of A: layout 101)
    opts = IROpts(precision=PREC[args.precision])
    rres = IRResult()

    def mixed_step():
        rc = zh.zmlhip_dgesv_mixed(101, n, 1, P(A), n, P(b), 1, P(x), 1, ctypes.byref(opts),
                                   ctypes.byref(rres), sh)
        if rc != 0:
            raise RuntimeError(f"zmlhip_dgesv_mixed: {rc} {_lib.last_error()}")

    for _ in range(args.warmup):
        mixed_step()
    zh.zmlhip_profile_enable(1)
    zh.zmlhip_profile_collect(PROF_GEMM_LOWP, *(ctypes.byref(v) for v in
                              (ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double())))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, gm, fb = [], [], []
    for _ in range(args.steps):
        mixed_step()
        iters.append(rres.iterations)
        gm.append(rres.gmres_iterations)
        fb.append(rres.fallback)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    zh.zmlhip_profile_enable(0)
    ms, nl, fl, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    zh.zmlhip_profile_collect(PROF_GEMM_LOWP, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(fl),
                              ctypes.byref(by))
    be_mixed = backward_error(x)
    err_true = ((x - x_true).abs().max() / x_true.abs().max()).item()
    flops = (2.0 / 3.0) * n ** 3
    valid = be_mixed <= 16.0

    # ---- plain FP64 on the same system
    Aw = torch.empty_like(A)
    ipiv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    x64 = torch.empty_like(b)

    def fp64_step():
        Aw.copy_(A)
        x64.copy_(b)
        rc = zh.zmlhip_dgetrf(101, n, n, P(Aw), n, P(ipiv), P(info), sh)
        rc |= zh.zmlhip_dgetrs(101, b"N", n, 1, P(Aw), n, P(ipiv), P(x64), 1, sh)
        if rc != 0:
            raise RuntimeError(f"FP64 path: {_lib.last_error()}")

    t_fp64 = be_fp64 = None
    if args.fp64_steps > 0:
        fp64_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            fp64_step()
        torch.cuda.synchronize()
        t_fp64 = (time.perf_counter() - t0) / args.fp64_steps
        be_fp64 = backward_error(x64)
    del Aw

    prec = args.precision
    roofline = roofline_of(f"mp_gemm ({prec} inputs, FP32 accumulate) trailing update", ms.value, nl.value,
                           fl.value, by.value, PEAK[prec], args.steps,
                           traffic_from_profiles(f"traffic_mp_gemm_{prec}_n{n}.json"))
    out = {
        "metric": METRIC,
        "value": round(flops / t_mixed / 1e12, 3) if valid else None,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{prec} update / fp32 panels / fp64 refinement",
        "data": "synthetic (device-generated A = diag(r)(R + delta n I)diag(c), b = A x_true)",
        "config": {"workload": f"dgesv mixed+IR, n={n}, delta={args.delta}, 1 RHS, row-major",
                   "n": n, "delta": args.delta, "precision": prec, "nrhs": 1},
        "backward_error": round(be_mixed, 4),
        "error_vs_x_true": err_true,
        "ir_iterations": iters, "gmres_iterations": gm, "fallback": fb,
        "plain_fp64": None if t_fp64 is None else {
            "ms": round(t_fp64 * 1e3, 2), "tflops": round(flops / t_fp64 / 1e12, 3),
            "backward_error": round(be_fp64, 4),
            "frac_fp64_peak": round(flops / t_fp64 / 1e12 / PEAK["fp64"], 4)},
        "speedup_vs_plain_fp64": None if t_fp64 is None else round(t_fp64 / t_mixed, 3),
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


def bench_plain(args):
    import torch
    from zml_amd import _lib
    zh = _lib.load()
    dev = torch.device("cuda")
    sh = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    n, ROW, COL = args.n, 101, 102
    g = torch.Generator(device=dev)
    g.manual_seed(42)
    A = torch.rand(n, n, dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    b = torch.rand(n, 1, dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    Aw = torch.empty_like(A)
    x = torch.empty_like(b)
    ipiv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)

    def be_of(M, xv, anorm):
        res = b - M @ xv
        return (res.abs().max() / (anorm * xv.abs().max() * n * 2.0 ** -53)).item()

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            step()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return ts[len(ts) // 2], ts

    def lu_step():
        Aw.copy_(A)
        x.copy_(b)
        info.zero_()
        rc = zh.zmlhip_dgetrf(ROW, n, n, P(Aw), n, P(ipiv), P(info), sh)
        rc |= zh.zmlhip_dgetrs(ROW, b"N", n, 1, P(Aw), n, P(ipiv), P(x), 1, sh)
        if rc != 0:
            raise RuntimeError(f"dgetrf/dgetrs: {_lib.last_error()}")

    zh.zmlhip_profile_enable(1)
    zh.zmlhip_profile_collect(PROF_GEMM_F64, *(ctypes.byref(v) for v in
                              (ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double())))
    for _ in range(args.warmup):
        lu_step()
    zh.zmlhip_profile_collect(PROF_GEMM_F64, *(ctypes.byref(v) for v in
                              (ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double())))
    t_lu, ts_lu = timed(lu_step, args.steps, 0)
    zh.zmlhip_profile_enable(0)
    ms, nl, fl, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    zh.zmlhip_profile_collect(PROF_GEMM_F64, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(fl),
                              ctypes.byref(by))
    assert int(info.item()) == 0
    be_lu = be_of(A, x, A.abs().sum(1).max().item())

    # The same solve through the host-pointer drop-in (LAPACKE_dgesv, row-major, pageable host
    # buffers): each call allocates its device buffers, copies A and b in over PCIe, factors,
    # solves and copies LU, ipiv and x back. Reported beside the device-resident figure, never
    # as the value.
    host_abi = None
    if not args.no_host_abi:
        import numpy as np
        a_host = A.cpu().numpy()
        b_host = b.cpu().numpy()
        x_dev = x.cpu().numpy()
        vp = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)

        def host_step():
            a_c, b_c, ip = a_host.copy(), b_host.copy(), np.zeros(n, np.int32)
            t0 = time.perf_counter()
            rc = zh.LAPACKE_dgesv(ROW, n, 1, vp(a_c), n, vp(ip), vp(b_c), 1)
            dt = time.perf_counter() - t0
            if rc != 0:
                raise RuntimeError(f"LAPACKE_dgesv returned {rc}: {_lib.last_error()}")
            return dt, b_c

        host_step()
        runs = [host_step() for _ in range(max(1, min(args.steps, 3)))]
        th = sorted(r[0] for r in runs)
        t_host = th[len(th) // 2]
        host_abi = {"entry": "LAPACKE_dgesv (host pointers, pageable memory)",
                    "ms": round(t_host * 1e3, 2), "tflops": round((2.0 / 3.0) * n ** 3 / t_host / 1e12, 3),
                    "pcie_bytes": int(8 * (2 * n * n + 2 * n) + 4 * n),
                    "ms_all": [round(t * 1e3, 2) for t in th],
                    "x_bitwise_equal_to_device_path": bool(np.array_equal(runs[-1][1].reshape(-1),
                                                                          x_dev.reshape(-1)))}
        del a_host

    # SPD: A = B B^T / n + I (B ~ N(0, 1)), column-major lower
    Bm = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    S = Bm @ Bm.T
    del Bm
    S.div_(n)
    S.diagonal().add_(1.0)
    S = 0.5 * (S + S.T)

    def chol_step():
        Aw.copy_(S)
        x.copy_(b)
        info.zero_()
        rc = zh.zmlhip_dpotrf(COL, b"L", n, P(Aw), n, P(info), sh)
        rc |= zh.zmlhip_dpotrs(COL, b"L", n, 1, P(Aw), n, P(x), n, sh)
        if rc != 0:
            raise RuntimeError(f"dpotrf/dpotrs: {_lib.last_error()}")

    t_ch, ts_ch = timed(chol_step, args.steps, args.warmup)
    assert int(info.item()) == 0
    be_ch = be_of(S, x, S.abs().sum(1).max().item())
    del S

    # dgemm n x n x n: C = A B (row-major), B = A^T's storage reused as a second operand
    Cm = torch.empty_like(A)

    def gemm_step():
        rc = zh.zmlhip_dgemm(ROW, 111, 111, n, n, n, 1.0, P(A), n, P(Aw), n, 0.0, P(Cm), n, sh)
        if rc != 0:
            raise RuntimeError(f"dgemm: {_lib.last_error()}")

    Aw.copy_(A.T)
    t_gm, ts_gm = timed(gemm_step, args.steps, args.warmup)
    ref = A[:64].to(torch.float64) @ Aw
    gemm_err = ((Cm[:64] - ref).abs().max() / ref.abs().max()).item()
    del Cm, Aw

    f_lu, f_ch, f_gm = (2.0 / 3.0) * n ** 3, n ** 3 / 3.0, 2.0 * n ** 3
    valid = be_lu <= 16 and be_ch <= 16
    rate = lambda f, t: round(f / t / 1e12, 3)
    out = {
        "metric": "plain FP64 LU solve TFLOP/s (dgetrf + dgetrs)",
        "value": rate(f_lu, t_lu) if valid else None,
        "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_lu * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (device-generated A ~ U[-1, 1]; SPD B B^T / n + I; 1 RHS)",
        "config": {"workload": f"plain FP64 dgetrf+dgetrs, dpotrf+dpotrs, dgemm, n={n}, row-major LU",
                   "n": n, "nrhs": 1},
        "timing": "median of the timed steps, each bracketed by device synchronisations",
        "lu": {"ms": round(t_lu * 1e3, 2), "tflops": rate(f_lu, t_lu), "backward_error": round(be_lu, 4),
               "frac_fp64_peak": round(f_lu / t_lu / 1e12 / PEAK["fp64"], 4),
               "ms_all": [round(t * 1e3, 2) for t in ts_lu]},
        "cholesky": {"ms": round(t_ch * 1e3, 2), "tflops": rate(f_ch, t_ch), "backward_error": round(be_ch, 4),
                     "frac_fp64_peak": round(f_ch / t_ch / 1e12 / PEAK["fp64"], 4),
                     "ms_all": [round(t * 1e3, 2) for t in ts_ch]},
        "dgemm": {"ms": round(t_gm * 1e3, 2), "tflops": rate(f_gm, t_gm), "rel_err_first_rows": gemm_err,
                  "frac_fp64_peak": round(f_gm / t_gm / 1e12 / PEAK["fp64"], 4)},
        "roofline": roofline_of("gemm_f64 trailing update inside dgetrf", ms.value, nl.value, fl.value, by.value,
                                PEAK["fp64"], args.steps, None),
        "host_abi": host_abi,
    }
    print(json.dumps(out))


def bench_cholesky(args):
    import torch
    from zml_amd import _lib
    from zml_amd._lib import IROpts, IRResult
    zh = _lib.load()
    dev = torch.device("cuda")
    sh = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    n, nrhs, COL = args.n, args.nrhs, 102
    g = torch.Generator(device=dev)
    g.manual_seed(44)
    Bm = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
    A = Bm @ Bm.T
    del Bm
    A.div_(n)
    A.diagonal().add_(args.mu)
    A = 0.5 * (A + A.T)
    # right-hand sides, column-major n x nrhs: stored as the rows of an nrhs x n tensor
    Bt = torch.rand(nrhs, n, dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    Xt = torch.zeros_like(Bt)
    torch.cuda.synchronize()
    anorm = A.abs().sum(1).max().item()

    def worst_be(X):
        R = Bt - X @ A          # rows of R are the residual columns (A symmetric)
        return max((R[q].abs().max() / (anorm * X[q].abs().max() * n * 2.0 ** -53)).item() for q in range(nrhs))

    opts = IROpts(precision=PREC[args.precision])
    res = IRResult()

    def mixed_step():
        rc = zh.zmlhip_dposv_mixed(COL, b"L", n, nrhs, P(A), n, P(Bt), n, P(Xt), n, ctypes.byref(opts),
                                   ctypes.byref(res), sh)
        if rc != 0:
            raise RuntimeError(f"zmlhip_dposv_mixed: {rc} {_lib.last_error()}")

    for _ in range(args.warmup):
        mixed_step()
    zh.zmlhip_profile_enable(1)
    zh.zmlhip_profile_collect(PROF_GEMM_LOWP, *(ctypes.byref(v) for v in
                              (ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double())))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it = []
    for _ in range(args.steps):
        mixed_step()
        it.append((res.iterations, res.gmres_iterations, res.fallback, res.info))
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    zh.zmlhip_profile_enable(0)
    ms, nl, fl, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    zh.zmlhip_profile_collect(PROF_GEMM_LOWP, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(fl),
                              ctypes.byref(by))
    be_mixed = worst_be(Xt)

    t_fp64 = be_fp64 = None
    if args.fp64_steps > 0:
        Aw = torch.empty_like(A)
        X64 = torch.empty_like(Bt)
        info = torch.zeros(1, dtype=torch.int32, device=dev)

        def fp64_step():
            Aw.copy_(A)
            X64.copy_(Bt)
            rc = zh.zmlhip_dpotrf(COL, b"L", n, P(Aw), n, P(info), sh)
            rc |= zh.zmlhip_dpotrs(COL, b"L", n, nrhs, P(Aw), n, P(X64), n, sh)
            if rc != 0:
                raise RuntimeError(f"FP64 Cholesky: {_lib.last_error()}")

        fp64_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            fp64_step()
        torch.cuda.synchronize()
        t_fp64 = (time.perf_counter() - t0) / args.fp64_steps
        be_fp64 = worst_be(X64)
        del Aw
    flops = n ** 3 / 3.0
    prec = args.precision
    out = {
        "metric": "Cholesky solve TFLOP/s at FP64 backward error, mixed+IR vs plain FP64",
        "value": round(flops / t_mixed / 1e12, 3) if be_mixed <= 16 else None,
        "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"{prec} update / fp32 L21 / fp64 diagonal blocks and refinement",
        "data": "synthetic (device-generated A = B B^T / n + mu I, B ~ N(0,1); 16 RHS ~ U[-1,1])",
        "config": {"workload": f"dposv mixed+IR, n={n}, mu={args.mu}, {nrhs} RHS, column-major lower",
                   "n": n, "mu": args.mu, "precision": prec, "nrhs": nrhs},
        "backward_error": round(be_mixed, 4),
        "ir_gmres_fallback_info": it,
        "plain_fp64": None if t_fp64 is None else {
            "ms": round(t_fp64 * 1e3, 2), "tflops": round(flops / t_fp64 / 1e12, 3),
            "backward_error": round(be_fp64, 4),
            "frac_fp64_peak": round(flops / t_fp64 / 1e12 / PEAK["fp64"], 4)},
        "speedup_vs_plain_fp64": None if t_fp64 is None else round(t_fp64 / t_mixed, 3),
        "roofline": roofline_of(f"mp_gemm ({prec}) lower-triangle trailing update", ms.value, nl.value, fl.value,
                                by.value, PEAK[prec], args.steps, None),
    }
    print(json.dumps(out))


def bench_batched(args):
    import numpy as np
    import torch
    from zml_amd import _lib
    from zml_amd._lib import IROpts
    zh = _lib.load()
    dev = torch.device("cuda")
    sh = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(47)
    count = args.systems
    ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(256), count))).astype(np.int64)
    nr = rng.integers(1, 5, count).astype(np.int64)
    spd = (np.arange(count) % 2 == 1)
    oa = np.concatenate([[0], np.cumsum(ns * ns)])
    ob = np.concatenate([[0], np.cumsum(ns * nr)])
    g = torch.Generator(device=dev)
    g.manual_seed(47)
    A = torch.rand(int(oa[-1]), dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    Bv = torch.rand(int(ob[-1]), dtype=torch.float64, device=dev, generator=g).mul_(2).sub_(1)
    X = torch.zeros_like(Bv)
    gen = np.nonzero(~spd)[0]
    didx = np.concatenate([oa[i] + np.arange(ns[i]) * (ns[i] + 1) for i in gen])
    dval = np.concatenate([np.full(ns[i], ns[i] / 4.0) for i in gen])
    A.index_add_(0, torch.from_numpy(didx).to(dev), torch.from_numpy(dval).to(dev))
    for n in np.unique(ns[spd]):
        ids = np.nonzero(spd & (ns == n))[0]
        G = torch.randn(len(ids), int(n), int(n), dtype=torch.float64, device=dev, generator=g)
        S = torch.bmm(G, G.transpose(1, 2).contiguous()).div_(float(n))
        S.diagonal(dim1=1, dim2=2).add_(1.0)
        idx = torch.from_numpy(np.concatenate([oa[i] + np.arange(n * n) for i in ids])).to(dev)
        A[idx] = S.reshape(-1)
    torch.cuda.synchronize()
    P = ctypes.c_void_p * count
    I = ctypes.c_int * count
    base_a, base_b, base_x = A.data_ptr(), Bv.data_ptr(), X.data_ptr()
    pa = P(*[base_a + 8 * int(o) for o in oa[:-1]])
    pb = P(*[base_b + 8 * int(o) for o in ob[:-1]])
    px = P(*[base_x + 8 * int(o) for o in ob[:-1]])
    plan = ctypes.c_void_p()
    t0 = time.perf_counter()
    rc = zh.zmlhip_batch_plan_create(ctypes.byref(plan), count, I(*ns.tolist()), I(*nr.tolist()),
                                     I(*spd.astype(int).tolist()), pa, pb, px, ctypes.byref(IROpts()), 1, sh)
    if rc != 0:
        raise RuntimeError(f"zmlhip_batch_plan_create: {rc} {_lib.last_error()}")
    t_plan = time.perf_counter() - t0
    for _ in range(args.warmup):
        zh.zmlhip_batch_plan_execute(plan, sh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        zh.zmlhip_batch_plan_execute(plan, sh)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    st = I()
    zh.zmlhip_batch_plan_status(plan, st, sh)
    status = np.array(st[:])
    tg, rr, cnt = (ctypes.c_int * 4)(), (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
    zh.zmlhip_batch_plan_shapes(plan, tg, rr, cnt)
    # backward error of a sample of systems
    hA, hB, hX = A.cpu().numpy(), Bv.cpu().numpy(), X.cpu().numpy()
    worst = 0.0
    for i in rng.choice(count, 300, replace=False):
        n, k = int(ns[i]), int(nr[i])
        Ai = hA[oa[i]:oa[i + 1]].reshape(n, n).T
        Bi = hB[ob[i]:ob[i + 1]].reshape(k, n).T
        Xi = hX[ob[i]:ob[i + 1]].reshape(k, n).T
        an = np.abs(Ai).sum(1).max()
        for q in range(k):
            worst = max(worst, np.abs(Bi[:, q] - Ai @ Xi[:, q]).max() / (an * np.abs(Xi[:, q]).max() * n * 2.0 ** -53))
    flops = float(np.sum(np.where(spd, ns ** 3 / 3.0, 2.0 * ns ** 3 / 3.0)))
    steps = np.maximum(status, 0)
    # A read for the inversion and once per residual (up to `steps` per right-hand side)
    bytes_ = float(np.sum(8.0 * ns * ns * (1 + nr * steps) + 16.0 * ns * nr))
    zh.zmlhip_batch_plan_destroy(plan)
    out = {
        "metric": "batched uneven solves, systems/s (TFLOP/s) at FP64 backward error",
        "value": round(count / t, 1) if (status >= 0).all() and worst <= 16 else None,
        "unit": "systems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 factors / fp64 refinement",
        "data": "synthetic (device-generated; n log-uniform [8,256], half U[-1,1]+nI/4, half BB^T/n+I, 1-4 RHS)",
        "config": {"workload": f"{count} independent systems, one hipGraph launch", "systems": count},
        "tflops": round(flops / t / 1e12, 3),
        "hbm": {"bytes_per_step": bytes_, "achieved_GBps": round(bytes_ / t / 1e9, 1), "peak_GBps": 8000,
                "frac": round(bytes_ / t / 8e12, 4)},
        "refinement_steps_max": int(status.max()), "not_converged": int((status < 0).sum()),
        "backward_error_sample_max": round(worst, 4),
        "bucket_shapes": [{"n_max": (32, 64, 128, 256)[b], "systems": cnt[b], "thread_rows": tg[b],
                           "matrix_rows_per_thread": rr[b]} for b in range(4)],
        "plan_seconds": round(t_plan, 3),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batched(hA, hB, ns, nr, spd, oa, ob)
    print(json.dumps(out))


def cpu_baseline_batched(hA, hB, ns, nr, spd, oa, ob, sample=20000):
    """zml's gesv (general) and posv (SPD) restated in C, one system per
    thread on every host core, over the first `sample` systems."""
    import concurrent.futures as cf
    import numpy as np
    import oracle as O
    O.lib()
    cores = len(os.sched_getaffinity(0))

    def solve(i):
        n, k = int(ns[i]), int(nr[i])
        a = hA[oa[i]:oa[i + 1]].copy()
        b = hB[ob[i]:ob[i + 1]].copy()
        if spd[i]:
            return O.posv(102, 122, n, k, a, n, b, n)
        return O.gesv(102, n, k, a, n, np.zeros(n, np.int32), b, n)

    idx = range(min(sample, len(ns)))
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        list(ex.map(solve, idx))
    t = time.perf_counter() - t0
    return {"value": round(len(idx) / t, 1), "unit": "systems/s", "cores": cores, "kind": "port",
            "sample": f"{'all' if len(idx) == len(ns) else 'first'} {len(idx)} systems of the same batch, zml gesv/posv restated in C "
                      f"(oracle/zml_oracle.c), one system per thread, {t:.2f} s"}


if __name__ == "__main__":
    main()
