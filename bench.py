#!/usr/bin/env python3
"""Benchmark of the level-3 hot path on one MI355X.

Headline workload (the metric "dgesv N=16384 time to FP64 backward error:
FP32-LU+IR vs FP64 LU; DGEMM TF/s"): A = Q1*diag(s)*Q2 of order 16384 in
FP64, Q1 and Q2 products of 8 Householder reflectors with Gaussian vectors,
s geometric from 1 to 1/kappa (kappa = 1e6 by default), b Gaussian, NRHS=1.
One step = one mixed-precision solve (dsgesv_: FP32 LU on MFMA + FP64
iterative refinement) on device-resident A and b.  The plain FP64 solve
(dgesv_) is timed in the same run, and DGEMM 8192^3 NN supplies the
roofline record.  The reference CPU path (the OpenBLAS library installed on
the host, called through its exported entry points) is timed beside it on
a bounded sample.

Prints ONE JSON line.  Data are synthetic (seeded).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import openblas_amd as oa  # noqa: E402

PEAK_FP64 = 78.6e12   # MI355X dense FP64 MFMA, published
PEAK_FP32 = 157.3e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(L, n, kappa, nrhs, seed):
    """A = Q1 diag(s) Q2 built on the device with rank-1 updates (16 x O(n^2)),
    b Gaussian.  Returns device buffers and host copies of the Householder
    vectors (enough to rebuild A on the host for the error report)."""
    rng = np.random.default_rng(seed)
    s = kappa ** (-np.arange(n) / (n - 1))
    A0 = np.zeros((n, n), order="F")
    A0[np.diag_indices(n)] = s
    dA = oa.DeviceBuffer.from_array(A0)
    del A0
    dv = oa.DeviceBuffer(8 * n)
    dw = oa.DeviceBuffer(8 * n)
    for side in ("left", "right"):
        for _ in range(8):
            v = rng.standard_normal(n)
            L.openblas_amd_memcpy(dv.ptr, v.ctypes.data, 8 * n, 0)
            c = -2.0 / float(v @ v)
            if side == "left":   # A <- A - c' v (A^T v)^T
                L.cblas_dgemm(102, 112, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
                L.cblas_dgemm(102, 111, 112, n, n, 1, c, dv.ptr, n, dw.ptr, n, 1.0, dA.ptr, n)
            else:                # A <- A - c' (A v) v^T
                L.cblas_dgemm(102, 111, 111, n, 1, n, 1.0, dA.ptr, n, dv.ptr, n, 0.0, dw.ptr, n)
                L.cblas_dgemm(102, 111, 112, n, n, 1, c, dw.ptr, n, dv.ptr, n, 1.0, dA.ptr, n)
    b = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dv.free()
    dw.free()
    return dA, b


def backward_error(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def time_dgemm(L, n, dt, reps, ta=111, tb=111):
    rng = np.random.default_rng(7)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(dt))
    dA, dB, dC = oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(A)
    f = L.cblas_dgemm if dt == np.float64 else L.cblas_sgemm
    f(102, ta, tb, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.5, dC.ptr, n)   # warm-up
    times = []
    for _ in range(reps):
        L.openblas_amd_timer_start()
        f(102, ta, tb, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.5, dC.ptr, n)
        times.append(L.openblas_amd_timer_stop())
    for d in (dA, dB, dC):
        d.free()
    ms = float(np.median(times))
    return ms, 2.0 * n ** 3 / (ms * 1e-3)


def cpu_baseline(A_host, b, budget_s=30.0):
    """The reference CPU path: dsgesv_ and dgesv_ of the installed OpenBLAS."""
    try:
        from oracle import refblas
        cfg = refblas.config()
    except Exception as e:  # no installed OpenBLAS: report absent
        return {"value": None, "unit": "solves/s", "cores": 0, "kind": "reference", "sample": f"absent: {e}"}
    cores = os.cpu_count() or 1
    refblas.set_num_threads(cores)
    A1 = np.array(A_host, order="F")
    t0 = time.perf_counter()
    X, _, it, info = refblas.dsgesv(A1, np.array(b, order="F"))
    t_ir = time.perf_counter() - t0
    res = {"value": 1.0 / t_ir, "unit": "solves/s", "cores": cores, "kind": "reference",
           "sample": f"one dsgesv_ N={A_host.shape[0]} NRHS={b.shape[1]} (ITER={it}) via {cfg}",
           "dsgesv_s": t_ir, "bwd_err": backward_error(A_host, X, b)}
    if t_ir < budget_s:
        A2 = np.array(A_host, order="F")
        b2 = np.array(b, order="F")
        t0 = time.perf_counter()
        refblas.gesv(A2, b2)
        res["dgesv_s"] = time.perf_counter() - t0
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    L = oa.lib()
    arch = ctypes.create_string_buffer(64)
    ncu, clk = ctypes.c_int(0), ctypes.c_int(0)
    L.openblas_amd_device_info(arch, 64, ctypes.byref(ncu), ctypes.byref(clk))
    n, nrhs = args.n, args.nrhs
    log(f"device {arch.value.decode()} CUs={ncu.value} clock={clk.value // 1000} MHz; building A (n={n}, kappa={args.kappa:g})")

    dA, b = make_system(L, n, args.kappa, nrhs, seed=2024)
    A_host = np.empty((n, n), order="F")
    dA.to_array(A_host)
    db = oa.DeviceBuffer.from_array(b)
    dX = oa.DeviceBuffer(8 * n * nrhs)
    dpiv = oa.DeviceBuffer(4 * n)
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    ip = oa.ip

    def ir_solve():
        L.dsgesv_(ip(n), ip(nrhs), dA.ptr, ip(n), dpiv.ptr, db.ptr, ip(n), dX.ptr, ip(n), None, None,
                  ctypes.byref(it), ctypes.byref(info))

    # --- FP32-LU + FP64 IR (the timed steps) ---
    for _ in range(args.warmup):
        ir_solve()
    L.openblas_amd_synchronize()
    ir_ms = []
    t_wall0 = time.perf_counter()
    for _ in range(args.steps):
        L.openblas_amd_timer_start()
        ir_solve()
        ir_ms.append(L.openblas_amd_timer_stop())
    L.openblas_amd_synchronize()
    t_wall = time.perf_counter() - t_wall0
    ir_iter, ir_info = it.value, info.value
    X = np.empty((n, nrhs), order="F")
    dX.to_array(X)
    if ir_iter < 0:   # fallback overwrote A with its FP64 factors: restore for the next solve
        dA.free()
        dA = oa.DeviceBuffer.from_array(A_host)
    bwd_ir = backward_error(A_host, X, b)
    log(f"dsgesv: {np.median(ir_ms):.1f} ms/solve ITER={ir_iter} info={ir_info} bwd={bwd_ir:.2e}")

    # --- plain FP64 LU solve on the same system ---
    dW = oa.DeviceBuffer(8 * n * n)
    dB2 = oa.DeviceBuffer(8 * n * nrhs)
    fp64_ms = []
    for rep in range(args.warmup + args.steps):
        L.openblas_amd_memcpy(dW.ptr, dA.ptr, 8 * n * n, 2)
        L.openblas_amd_memcpy(dB2.ptr, db.ptr, 8 * n * nrhs, 2)
        L.openblas_amd_timer_start()
        L.dgesv_(ip(n), ip(nrhs), dW.ptr, ip(n), dpiv.ptr, dB2.ptr, ip(n), ctypes.byref(info))
        t = L.openblas_amd_timer_stop()
        if rep >= args.warmup:
            fp64_ms.append(t)
    X64 = np.empty((n, nrhs), order="F")
    dB2.to_array(X64)
    bwd_64 = backward_error(A_host, X64, b)
    log(f"dgesv: {np.median(fp64_ms):.1f} ms/solve info={info.value} bwd={bwd_64:.2e}")
    for d in (dW, dB2):
        d.free()

    # --- DGEMM roofline record (8192^3 NN, device-resident) and 1024^3 ---
    g_ms, g_flops = time_dgemm(L, 8192, np.float64, 5)
    g1_ms, g1_flops = time_dgemm(L, 1024, np.float64, 20)
    s_ms, s_flops = time_dgemm(L, 8192, np.float32, 5)
    log(f"dgemm 8192: {g_ms:.2f} ms {g_flops / 1e12:.1f} TF/s; dgemm 1024: {g1_flops / 1e12:.1f} TF/s; "
        f"sgemm 8192: {s_flops / 1e12:.1f} TF/s")

    cpu = None
    if not args.no_cpu:
        log("timing the reference CPU path ...")
        cpu = cpu_baseline(A_host, b)
        log(f"cpu: {cpu}")

    ir_med = float(np.median(ir_ms))
    f64_med = float(np.median(fp64_ms))
    flops_solve = 2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs
    out = {
        "metric": "dgesv N=16384 time to FP64 backward error: FP32-LU+IR vs FP64 LU; DGEMM TF/s",
        "value": 1000.0 / ir_med,
        "unit": "solves/s (FP32-LU+FP64-IR dgesv, N=16384, NRHS=1, to FP64 backward error)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ir_med,
        "wall_s_timed_region": t_wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 LU + fp64 refinement",
        "data": "synthetic: A=Q1*diag(s)*Q2 (8+8 Householder reflectors, s geometric 1..1/kappa), b~N(0,1), seeded",
        "config": {"workload": f"dsgesv vs dgesv N={n} NRHS={nrhs} kappa={args.kappa:g}; dgemm 8192^3 NN",
                   "n": n, "nrhs": nrhs, "kappa": args.kappa},
        "ir": {"ms": ir_med, "iter": ir_iter, "info": ir_info, "backward_error": bwd_ir,
               "effective_tflops": flops_solve / (ir_med * 1e-3) / 1e12},
        "fp64_lu": {"ms": f64_med, "backward_error": bwd_64,
                    "tflops": flops_solve / (f64_med * 1e-3) / 1e12},
        "speedup_ir_vs_fp64": f64_med / ir_med,
        "dgemm_8192_tflops": g_flops / 1e12,
        "dgemm_1024_tflops": g1_flops / 1e12,
        "sgemm_8192_tflops": s_flops / 1e12,
        "roofline": {"kernel": "gemm_kernel<double,128,128,16> (dgemm 8192^3 NN)", "bound": "mfma",
                     "achieved": g_flops / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": g_flops / PEAK_FP64, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for d in (dA, db, dX, dpiv):
        d.free()


if __name__ == "__main__":
    main()
