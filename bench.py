#!/usr/bin/env python3
"""Benchmark of the HAMLSS solve path on one MI355X.

Workload (default): 2-D Poisson 5-point on a 2048 x 2048 grid (n = 4,194,304,
nnz = 20,963,328), b = A x*, x* ~ U[-1,1] (splitmix64, seed 42), PCG +
Jacobi to rel-res <= 1e-8.  A "step" is one full solve of that system with
inputs resident in HBM.

Reported (one JSON line):
  value         solves/s of the mixed-precision solve (FP32 SpMV / Krylov
                vectors, FP64 reductions, FP64 reliable-update refinement)
  fp64_baseline solves/s of the plain FP64 device PCG on the same system
  roofline      the CSR SpMV kernel (FP64 values and vectors) timed with HIP
                events on the context's stream; algorithmic bytes
                B = nnz*(8+4) + 4*(n+1) + 8*n + 8*n  (flop_model::spmv_bytes)
  cpu_baseline  the FP64 OpenMP CPU oracle (a port of the reference path) on a
                bounded sample: 200 PCG iterations at the full size, scaled by
                the iteration count the solve needs.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("hsps_amd", os.path.join(ROOT, "hybrid-sparse-pde-solver_amd",
                                                                           "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hsps_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class HipEvents:
    """hipEventRecord / hipEventElapsedTime on a given stream (libamdhip64)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def spmv_bytes(n, nnz, sv=8, sx=8):
    return nnz * (sv + 4) + 4 * (n + 1) + sx * n + sx * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    H = load_pkg()
    ev = HipEvents()
    ctx = H.Context(0)
    stream = H.lib().hsps_ctx_stream(ctx.h)
    side = args.grid
    n = side * side
    rp, ci, v = H.poisson_2d(side)
    nnz = len(v)
    A = H.Csr(ctx, n, n, rp, ci, v)
    xs = H.uniform_pm1(42, n)
    x_dev = ctx.array(xs)
    b_dev = ctx.array(n=n)
    A.spmv(x_dev, b_dev)  # b = A x* on the device (FP64)
    ctx.sync()
    b_host = b_dev.download()

    # ---- SpMV roofline (FP64 values and vectors) -------------------------------
    y_dev = ctx.array(n=n)
    for _ in range(5):
        A.spmv(x_dev, y_dev)
    e0, e1 = ev.event(), ev.event()
    ev.hip.hipEventRecord(e0, stream)
    for _ in range(args.spmv_reps):
        A.spmv(x_dev, y_dev)
    ev.hip.hipEventRecord(e1, stream)
    spmv_ms = ev.elapsed_ms(e0, e1) / args.spmv_reps
    B = spmv_bytes(n, nnz)
    spmv_gbs = B / (spmv_ms * 1e-3) / 1e9
    # FP32 SpMV
    x32, y32 = ctx.array(xs.astype(np.float32), dtype=np.float32), ctx.array(n=n, dtype=np.float32)
    for _ in range(5):
        A.spmv(x32, y32, precision=H.HSPS_MIXED)
    ev.hip.hipEventRecord(e0, stream)
    for _ in range(args.spmv_reps):
        A.spmv(x32, y32, precision=H.HSPS_MIXED)
    ev.hip.hipEventRecord(e1, stream)
    spmv32_ms = ev.elapsed_ms(e0, e1) / args.spmv_reps
    B32 = spmv_bytes(n, nnz, 4, 4)

    # ---- solves ------------------------------------------------------------------------
    def run(precision, steps, warmup):
        p = H.default_params(tol=1e-8, max_iter=20000, precision=precision, check_every=64)
        M = H.Precond(A, "jacobi", p)
        xd = ctx.array(n=n)
        for _ in range(warmup):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            H.solve(A, b_dev, xd, "cg", M, p)
        ctx.sync()
        times, stats = [], []
        for _ in range(steps):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            ctx.sync()
            t0 = time.perf_counter()
            ok, st = H.solve(A, b_dev, xd, "cg", M, p)
            ctx.sync()
            times.append(time.perf_counter() - t0)
            assert ok and st.true_residual <= 1e-8, (precision, st.true_residual)
            stats.append(st)
        x = xd.download()
        err = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
        return times, stats, err

    t64, s64, err64 = run("fp64", max(1, args.steps), args.warmup)
    tmx, smx, errmx = run("mixed", args.steps, args.warmup)
    total = sum(tmx)
    value = args.steps / total
    fp64_rate = len(t64) / sum(t64)

    # ---- CPU baseline: FP64 OpenMP oracle, bounded sample -----------------------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        Ao = O.Csr.from_csr(n, n, rp, ci, v)
        sample_iters = 200
        t0 = time.perf_counter()
        _, sto = O.solve(Ao, b_host, "cg", "jacobi", tol=1e-8, max_iter=sample_iters)
        dt = time.perf_counter() - t0
        per_iter = dt / max(sto.iterations, 1)
        need = s64[0].iterations
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": 1.0 / (per_iter * need), "unit": "solves/s", "cores": cores, "kind": "port",
               "sample": f"{sto.iterations} PCG+Jacobi iterations of the FP64 OpenMP oracle on the full "
                         f"{side}x{side} system ({dt:.1f} s), scaled to the {need} iterations a solve needs"}

    line = {
        "metric": "solves/s @ rel-res<=1e-8 (mixed vs plain FP64), p50/p99 latency",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded splitmix64 x*, b = A x*)",
        "config": {"workload": f"poisson2d_{side}x{side}_pcg_jacobi", "n": n, "nnz": nnz, "tol": 1e-8,
                   "solver": "PCG+Jacobi, mixed FP32/FP64 reliable updates"},
        "fp64_baseline": {"value": fp64_rate, "unit": "solves/s", "iterations": s64[0].iterations,
                          "ms_per_solve": 1e3 * sum(t64) / len(t64), "rel_err_vs_xstar": err64},
        "mixed": {"iterations": smx[0].iterations, "reliable_updates": smx[0].reliable_updates,
                  "true_residual": smx[0].true_residual, "rel_err_vs_xstar": errmx,
                  "p50_ms": 1e3 * float(np.percentile(tmx, 50)), "p99_ms": 1e3 * float(np.percentile(tmx, 99))},
        "speedup_vs_fp64": value / fp64_rate,
        "roofline": {"bound": "hbm", "kernel": "k_spmv<double,double,EpiStore> (CSR row-block SpMV, FP64)",
                     "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": spmv_gbs / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": B, "avg_launch_us": spmv_ms * 1e3,
                     "fp32_spmv": {"avg_launch_us": spmv32_ms * 1e3, "algorithmic_bytes": B32,
                                   "achieved": B32 / (spmv32_ms * 1e-3) / 1e9}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
