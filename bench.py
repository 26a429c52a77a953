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
PEAK_INT8 = 5.0e15
PEAK_HBM = 8.0e12


def profiled_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated with a copy of known size)."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


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


# ---- INT8-slice emulated DGEMM (M=N=K=8192, entries u*exp(phi*g)) -----------------
def bench_ozaki(L, n=8192, phis=(0.5, 2.0, 8.0), reps=3, cpu=True):
    res = []
    for phi in phis:
        rng = np.random.default_rng(int(phi * 100) + 11)
        A = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(phi * rng.standard_normal((n, n))))
        B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) * np.exp(phi * rng.standard_normal((n, n))))
        dA, dB = oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(B)
        dC, dC64 = oa.DeviceBuffer(8 * n * n), oa.DeviceBuffer(8 * n * n)
        L.cblas_dgemm(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC64.ptr, n)
        t64 = []
        for _ in range(reps):
            L.openblas_amd_timer_start()
            L.cblas_dgemm(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC64.ptr, n)
            t64.append(L.openblas_amd_timer_stop())
        s = L.openblas_amd_dgemm_ozaki(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 0)
        tz = []
        for _ in range(reps):
            L.openblas_amd_timer_start()
            L.openblas_amd_dgemm_ozaki(102, 111, 111, n, n, n, 1.0, dA.ptr, n, dB.ptr, n, 0.0, dC.ptr, n, 0)
            tz.append(L.openblas_amd_timer_stop())
        C = np.empty((n, n), order="F"); C64 = np.empty((n, n), order="F")
        dC.to_array(C); dC64.to_array(C64)
        rows = rng.choice(n, 32, replace=False)
        ref = A[rows] @ B
        gauge = np.abs(A[rows]) @ np.abs(B)
        err_oz = float(np.max(np.abs(C[rows] - ref) / gauge))
        err_64 = float(np.max(np.abs(C64[rows] - ref) / gauge))
        ms_z, ms_64 = float(np.median(tz)), float(np.median(t64))
        fl = 2.0 * n ** 3
        pairs = s * (s + 1) // 2
        r = {"phi": phi, "slices": s, "ms": ms_z, "fp64_equiv_tflops": fl / ms_z / 1e9,
             "frac_of_fp64_peak": fl / (ms_z * 1e-3) / PEAK_FP64,
             "int8_tops": pairs * fl / ms_z / 1e9, "frac_of_int8_peak": pairs * fl / (ms_z * 1e-3) / PEAK_INT8,
             "native_dgemm_ms": ms_64, "speedup_vs_native": ms_64 / ms_z,
             "err_vs_ref_normwise": err_oz, "native_err_vs_ref_normwise": err_64}
        if cpu:
            try:
                from oracle import refblas
                refblas.set_num_threads(os.cpu_count() or 1)
                Cs = np.zeros((1024, n), order="F")
                As = np.asfortranarray(A[:1024])
                t0 = time.perf_counter()
                refblas.cblas_gemm(False, False, 1.0, As, B, 0.0, Cs)
                r["cpu_dgemm_1024rows_s"] = time.perf_counter() - t0
                r["cpu_dgemm_tflops"] = 2.0 * 1024 * n * n / r["cpu_dgemm_1024rows_s"] / 1e12
            except Exception as e:
                r["cpu"] = f"absent: {e}"
        log(f"ozaki phi={phi}: s={s} {ms_z:.1f} ms ({r['fp64_equiv_tflops']:.1f} TF/s eq, native {ms_64:.1f} ms) "
            f"err={err_oz:.2e} native_err={err_64:.2e}")
        res.append(r)
        for d in (dA, dB, dC, dC64):
            d.free()
        del A, B, C, C64
    return res


# ---- batched small problems (10,000 problems, n ~ U{8..256}) ----------------------
def _ptrs(base, off, es):
    return (ctypes.c_void_p * len(off))(*[base + int(o) * es for o in off])


def _i32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def bench_batched(L, count=10000, seed=5, reps=3, cpu=True, loop_sample=300, cpu_sample=1000):
    rng = np.random.default_rng(seed)
    ns = rng.integers(8, 257, count).astype(np.int64)
    sq = ns * ns
    off = np.concatenate([[0], np.cumsum(sq)[:-1]])
    voff = np.concatenate([[0], np.cumsum(ns)[:-1]])
    total, vtotal = int(sq.sum()), int(ns.sum())
    n32 = _i32(ns)
    gemm_flops = float((2 * ns ** 3).sum())
    lu_flops = float((2.0 / 3.0 * ns ** 3).sum())
    rs_flops = float((2 * ns ** 2).sum())
    po_flops = float((ns ** 3 / 3.0).sum())
    out = {"count": count, "n_dist": "U{8..256}", "seed": seed}
    for tag, dt in (("d", np.float64), ("s", np.float32)):
        es = np.dtype(dt).itemsize
        r = {}
        A = rng.standard_normal(total).astype(dt)
        Bm = rng.standard_normal(total).astype(dt)
        dA, dB, dC, dW = (oa.DeviceBuffer.from_array(A), oa.DeviceBuffer.from_array(Bm), oa.DeviceBuffer(es * total),
                          oa.DeviceBuffer(es * total))
        pA, pB, pC, pW = (_ptrs(d.ptr, off, es) for d in (dA, dB, dC, dW))
        ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
        gemm_b = L.cblas_dgemm_batch if tag == "d" else L.cblas_sgemm_batch
        one = (ct * count)(*([1.0] * count)); zero = (ct * count)(*([0.0] * count))
        tn = _i32(np.full(count, 111)); tt = _i32(np.full(count, 112)); gs = _i32(np.ones(count))

        def timed(fn, prep=None):
            fn()   # first call: tuning + graph capture
            ts, ws = [], []
            for _ in range(reps):
                if prep:
                    prep()
                L.openblas_amd_synchronize()
                w0 = time.perf_counter()
                L.openblas_amd_timer_start()
                fn()
                ts.append(L.openblas_amd_timer_stop())
                ws.append((time.perf_counter() - w0) * 1e3)
            return float(np.median(ts)), float(np.median(ws))

        # (i) GEMM batch: group_count = count, group_size 1, C = A*B
        g = lambda: gemm_b(102, _pi(tn), _pi(tn), _pi(n32), _pi(n32), _pi(n32), one, pA, _pi(n32), pB, _pi(n32), zero,
                           pC, _pi(n32), count, _pi(gs))
        ms, wall = timed(g)
        byt = float((3 * sq).sum()) * es
        r["gemm"] = {"ms": ms, "wall_ms": wall, "tflops": gemm_flops / ms / 1e9,
                     "frac_of_peak": gemm_flops / (ms * 1e-3) / (PEAK_FP64 if tag == "d" else PEAK_FP32),
                     "hbm_frac": byt / (ms * 1e-3) / PEAK_HBM}
        # per-problem launch loop on a sample
        f1 = L.cblas_dgemm if tag == "d" else L.cblas_sgemm
        L.openblas_amd_synchronize()
        w0 = time.perf_counter()
        for i in range(loop_sample):
            n = int(ns[i])
            f1(102, 111, 111, n, n, n, 1.0, pA[i], n, pB[i], n, 0.0, pC[i], n)
        L.openblas_amd_synchronize()
        r["gemm"]["loop_ms_est"] = (time.perf_counter() - w0) * 1e3 * count / loop_sample

        # (ii) getrf + getrs batch (A ~ N(0,1), b ~ N(0,1), 1 RHS)
        dpiv = oa.DeviceBuffer(4 * vtotal)
        ppiv = _ptrs(dpiv.ptr, voff, 4)
        dbv, dx = oa.DeviceBuffer.from_array(rng.standard_normal(vtotal).astype(dt)), oa.DeviceBuffer(es * vtotal)
        pX = _ptrs(dx.ptr, voff, es)
        info = np.zeros(count, np.int32)
        fr = L.openblas_amd_dgetrf_batch if tag == "d" else L.openblas_amd_sgetrf_batch
        fs = L.openblas_amd_dgetrs_batch if tag == "d" else L.openblas_amd_sgetrs_batch
        reset_a = lambda: L.openblas_amd_memcpy(dW.ptr, dA.ptr, es * total, 2)
        ms, wall = timed(lambda: fr(_pi(n32), _pi(n32), pW, _pi(n32), ppiv, _pi(info), count), reset_a)
        r["getrf"] = {"ms": ms, "wall_ms": wall, "tflops": lu_flops / ms / 1e9,
                      "frac_of_peak": lu_flops / (ms * 1e-3) / (PEAK_FP64 if tag == "d" else PEAK_FP32),
                      "hbm_frac": float((2 * sq).sum()) * es / (ms * 1e-3) / PEAK_HBM, "info_nonzero": int((info != 0).sum())}
        reset_b = lambda: L.openblas_amd_memcpy(dx.ptr, dbv.ptr, es * vtotal, 2)
        one_rhs = _i32(np.ones(count))
        ms, wall = timed(lambda: fs(b"N", _pi(n32), _pi(one_rhs), pW, _pi(n32), ppiv, pX, _pi(n32), _pi(info), count),
                         reset_b)
        r["getrs"] = {"ms": ms, "wall_ms": wall, "gflops": rs_flops / ms / 1e6,
                      "hbm_frac": float(sq.sum()) * es / (ms * 1e-3) / PEAK_HBM}
        # residual spot check on 16 problems
        Xh = np.empty(vtotal, dt); dx.to_array(Xh)
        bh = np.empty(vtotal, dt); dbv.to_array(bh)
        worst = 0.0
        for i in rng.choice(count, 16, replace=False):
            n, o, v = int(ns[i]), int(off[i]), int(voff[i])
            Ai = A[o:o + n * n].reshape(n, n, order="F").astype(np.float64)
            xi, bi = Xh[v:v + n].astype(np.float64), bh[v:v + n].astype(np.float64)
            worst = max(worst, float(np.abs(bi - Ai @ xi).max() / (np.abs(Ai).sum(1).max() * np.abs(xi).max())))
        r["getrs"]["worst_backward_error_16"] = worst
        # per-problem loop (dgetrf_ on device pointers) on a sample
        f1 = L.dgetrf_ if tag == "d" else L.sgetrf_
        reset_a()
        L.openblas_amd_synchronize()
        ii = ctypes.c_int(0)
        w0 = time.perf_counter()
        for i in range(loop_sample):
            n = int(ns[i])
            f1(oa.ip(n), oa.ip(n), pW[i], oa.ip(n), ppiv[i], ctypes.byref(ii))
        r["getrf"]["loop_ms_est"] = (time.perf_counter() - w0) * 1e3 * count / loop_sample

        # (iii) potrf batch on A = B*B^T + n*I (built with the batched GEMM)
        Ch = np.zeros(total, dt)
        for i in range(count):
            n, o = int(ns[i]), int(off[i])
            Ch[o:o + n * n:n + 1] = n
        dS = oa.DeviceBuffer.from_array(Ch)
        del Ch
        pS = _ptrs(dS.ptr, off, es)
        gemm_b(102, _pi(tn), _pi(tt), _pi(n32), _pi(n32), _pi(n32), one, pB, _pi(n32), pB, _pi(n32), one, pS,
               _pi(n32), count, _pi(gs))
        fp = L.openblas_amd_dpotrf_batch if tag == "d" else L.openblas_amd_spotrf_batch
        reset_s = lambda: L.openblas_amd_memcpy(dW.ptr, dS.ptr, es * total, 2)
        ms, wall = timed(lambda: fp(b"L", _pi(n32), pW, _pi(n32), _pi(info), count), reset_s)
        r["potrf"] = {"ms": ms, "wall_ms": wall, "tflops": po_flops / ms / 1e9,
                      "frac_of_peak": po_flops / (ms * 1e-3) / (PEAK_FP64 if tag == "d" else PEAK_FP32),
                      "info_nonzero": int((info != 0).sum())}

        if cpu and tag == "d":
            try:
                from oracle import refblas
                cores = os.cpu_count() or 1
                refblas.set_num_threads(cores)
                sel = range(cpu_sample)
                As = [A[off[i]:off[i] + sq[i]].reshape(ns[i], ns[i], order="F").copy(order="F") for i in sel]
                Bs = [Bm[off[i]:off[i] + sq[i]].reshape(ns[i], ns[i], order="F").copy(order="F") for i in sel]
                Cs = [np.zeros((ns[i], ns[i]), order="F") for i in sel]
                # a loop of cblas_dgemm: the installed library's cblas_dgemm_batch
                # crashes on problems with m*n*k <= 100^3 (its small-matrix path)
                t0 = time.perf_counter()
                for a, b_, c in zip(As, Bs, Cs):
                    refblas.cblas_gemm(False, False, 1.0, a, b_, 0.0, c)
                tg = time.perf_counter() - t0
                t0 = time.perf_counter()
                for a in As:
                    pv, _ = refblas.getrf(a)
                    refblas.getrs(False, a, pv, np.ones((a.shape[0], 1), order="F"))
                tl = time.perf_counter() - t0
                Ss = [np.asfortranarray(b @ b.T + b.shape[0] * np.eye(b.shape[0])) for b in Bs]
                t0 = time.perf_counter()
                for a in Ss:
                    refblas.potrf(True, a)
                tp = time.perf_counter() - t0
                scale = count / cpu_sample
                r["cpu"] = {"cores": cores, "kind": "reference", "sample": f"first {cpu_sample} problems, scaled",
                            "gemm_loop_ms_est": tg * 1e3 * scale, "getrf_getrs_loop_ms_est": tl * 1e3 * scale,
                            "potrf_loop_ms_est": tp * 1e3 * scale}
            except Exception as e:
                r["cpu"] = f"absent: {e}"
        log(f"batched {tag}: " + json.dumps({k: (v if not isinstance(v, dict) else {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}) for k, v in r.items()}))
        out[tag] = r
        for d in (dA, dB, dC, dW, dpiv, dbv, dx, dS):
            d.free()
        del A, Bm
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ozaki", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--section", default="all", choices=["all", "solve", "dgemm", "ozaki", "batched"],
                    help="run one workload only (for profiling); prints a partial JSON line")
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
    if args.section == "dgemm":
        g_ms, g_flops = time_dgemm(L, 8192, np.float64, 5)
        # a device-to-device copy of known size: the FETCH_SIZE / WRITE_SIZE
        # calibration for the counter runs of this section
        nb = 512 << 20
        c1, c2 = oa.DeviceBuffer(nb), oa.DeviceBuffer(nb)
        L.openblas_amd_memcpy(c2.ptr, c1.ptr, nb, 2)
        L.openblas_amd_synchronize()
        print(json.dumps({"section": "dgemm", "dgemm_8192_ms": g_ms, "tflops": g_flops / 1e12,
                          "calibration_copy_bytes": nb}), flush=True)
        return
    if args.section == "ozaki":
        print(json.dumps({"section": "ozaki", "ozaki": bench_ozaki(L, cpu=False, reps=2)}), flush=True)
        return
    if args.section == "batched":
        print(json.dumps({"section": "batched", "batched": bench_batched(L, cpu=False, reps=2, loop_sample=50)}),
              flush=True)
        return
    if args.section == "solve":
        args.no_ozaki = args.no_batched = args.no_cpu = True
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
    g_ms = g_flops = g1_flops = s_flops = float("nan")
    if args.section == "all":
        g_ms, g_flops = time_dgemm(L, 8192, np.float64, 5)
        g1_ms, g1_flops = time_dgemm(L, 1024, np.float64, 20)
        s_ms, s_flops = time_dgemm(L, 8192, np.float32, 5)
        log(f"dgemm 8192: {g_ms:.2f} ms {g_flops / 1e12:.1f} TF/s; dgemm 1024: {g1_flops / 1e12:.1f} TF/s; "
            f"sgemm 8192: {s_flops / 1e12:.1f} TF/s")

    ozaki = None if args.no_ozaki else bench_ozaki(L, cpu=not args.no_cpu)
    batched = None if args.no_batched else bench_batched(L, cpu=not args.no_cpu)

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
        "roofline": {"kernel": "gemm_kernel<double,128,128,4> (dgemm 8192^3 NN)", "bound": "mfma",
                     "achieved": g_flops / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": g_flops / PEAK_FP64, "traffic": profiled_traffic("dgemm_8192_nn_fp64.md"),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)"},
        "cpu_baseline": cpu,
        "ozaki_dgemm_8192": ozaki,
        "batched_10000": batched,
    }
    print(json.dumps(out), flush=True)
    for d in (dA, db, dX, dpiv):
        d.free()


if __name__ == "__main__":
    main()
