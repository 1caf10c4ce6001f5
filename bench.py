#!/usr/bin/env python3
"""Benchmark of the HAMLSS solve path on one MI355X.

Default workload ("batched", 512 systems): grid side log-uniform in
[32, 512], half 2-D Poisson (5-point, unscaled) and half anisotropic
diffusion -eps u_xx - u_yy (eps log-uniform in [1e-3, 1]), seeded; b = A x*,
x* ~ U[-1,1] (splitmix64).  Every system is solved to FP64 true rel-res
<= 1e-8 with PCG + Jacobi.  A "step" is one solve of all 512 systems with
inputs resident in HBM.

  value          solves/s of the batched mixed-precision solve (FP32 SpMV /
                 Krylov vectors, FP64 reductions, FP64 reliable-update
                 refinement), all 512 systems in one batch
  fp64_baseline  solves/s of the plain FP64 device PCG, one system at a time
  roofline       the batched iteration at full load (K1 + K2, each with its
                 per-system finish fused in, over every tiled row): 34 B of
                 algorithmic traffic per row (stencil storage, FP32), timed as
                 (t(96 iterations) - t(32 iterations)) / 64
  spmv_roofline  the CSR SpMV kernel (FP64) on the 2048^2 Poisson matrix
                 (n = 4,194,304; a working set above the 256 MB Infinity
                 Cache), timed with HIP events on the context's stream;
                 algorithmic bytes B = nnz*(8+4) + 4*(n+1) + 8*n + 8*n
  cpu_baseline   the FP64 OpenMP CPU oracle (port of the reference path) on a
                 bounded sample of the same systems, scaled by work

--workload single runs the 2048^2 Poisson system alone (PCG+Jacobi, mixed vs
FP64) instead.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# The serving leg runs one context (a stream of its own) per worker plus the
# batch thread's.  HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware
# queues (4 by default), and kernels of streams that share a queue run one
# after another, so a small request would wait behind a large one: give the
# streams queues of their own.  Read when HIP initialises, i.e. before the
# first device call below; a value set by the caller is kept.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of the FP64 SpMV on the 2048^2 Poisson matrix, from
# two rocprofv3 --pmc runs of `bench.py --spmv-only` (FETCH_SIZE 154,831 KiB
# x 2.000 + WRITE_SIZE 32,768 KiB x 1.000; the factors make a 256 MiB copy
# read and write 256 MiB).
SPMV_FP64_HBM_BYTES = (154831 * 2.0 + 32768 * 1.0) * 1024
# HBM bytes per row-iteration of the stencil-stored FP32 batched iteration
# (k_b_k1v + k_b_k2v), from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
# `bench.py --full-load-only --tile-rows 2048`: the median full-load dispatch
# of each kernel (factors 2.000 / 1.000 from the 256 MiB copy of the same run)
# divided by the 21,108,314 tiled rows (profiles/batched512_pcg_jacobi_kernels.md).
BATCHED_HBM_BYTES_PER_ROW_ITER = 35.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def spmv_bytes(n, nnz, sv=8, sx=8):
    return nnz * (sv + 4) + 4 * (n + 1) + sx * n + sx * n


def physical_cores():
    """Physical cores among the CPUs this process may run on (lscpu core /
    socket ids of the affinity set), capped by OMP_NUM_THREADS when it is set
    (the share of a shared host)."""
    import subprocess
    allowed = os.sched_getaffinity(0)
    cores = set()
    try:
        txt = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        for line in txt.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(t) for t in line.split(",")[:3])
            if cpu in allowed:
                cores.add((sock, core))
    except (OSError, subprocess.CalledProcessError, ValueError):
        cores = set(allowed)
    n = max(1, len(cores))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return n, len(allowed)


def cpu_oracle_runs(specs, b, row_off, sample, threads, warmup, reps, tag, want_x=False):
    """Times the FP64 OpenMP oracle (oracle/cpu_baseline.py, a child process
    with its own OpenMP runtime) on the sample systems at `threads` threads
    (OMP_PROC_BIND=close, OMP_PLACES=cores): `warmup` untimed passes, then the
    median of `reps` timed ones."""
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="hsps_cpu_")
    inp, out, xs = (os.path.join(tmp, f) for f in ("in.npz", "out.json", "x.npz"))
    kinds = np.array([0 if specs[s][0] == "poisson" else 1 for s in sample], np.int32)
    sides = np.array([specs[s][1] for s in sample], np.int32)
    eps = np.array([specs[s][2] for s in sample], np.float64)
    bb = [b[row_off[s]:row_off[s + 1]] for s in sample]
    off = np.zeros(len(sample) + 1, np.int64)
    np.cumsum([len(v) for v in bb], out=off[1:])
    np.savez(inp, kinds=kinds, sides=sides, eps=eps, b=np.concatenate(bb), b_off=off, warmup=warmup, reps=reps)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), inp, out] + (["--solutions", xs] if want_x else [])
    log(f"[bench] cpu baseline ({tag}): {len(sample)} systems, {threads} threads, {warmup} warm-up + {reps} timed passes")
    subprocess.run(cmd, env=env, check=True)
    with open(out) as f:
        res = json.load(f)
    x = None
    if want_x:
        x = np.load(xs)["x"]
        res["x_off"] = off
    for path in (inp, out, xs):
        if os.path.exists(path):
            os.remove(path)
    os.rmdir(tmp)
    return res, x


def spmv_roofline(H, ctx, ev, side=2048, reps=50):
    stream = H.lib().hsps_ctx_stream(ctx.h)
    n = side * side
    rp, ci, v = H.poisson_2d(side)
    A = H.Csr(ctx, n, n, rp, ci, v)
    x = ctx.array(H.uniform_pm1(42, n))
    y = ctx.array(n=n)
    out = {}
    for prec, name, sv in [(H.HSPS_FP64, "fp64", 8), (H.HSPS_MIXED, "fp32", 4)]:
        xx = x if sv == 8 else ctx.array(H.uniform_pm1(42, n).astype(np.float32), dtype=np.float32)
        yy = y if sv == 8 else ctx.array(n=n, dtype=np.float32)
        for _ in range(5):
            A.spmv(xx, yy, precision=prec)
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, stream)
        for _ in range(reps):
            A.spmv(xx, yy, precision=prec)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1) / reps
        B = spmv_bytes(n, len(v), sv, sv)
        out[name] = {"avg_launch_us": ms * 1e3, "algorithmic_bytes": B, "achieved_GBs": B / (ms * 1e-3) / 1e9}
    A.close()
    return out, n, len(v)


def c4_specs(count=512, seed=2024):
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), count))).astype(int)
    kinds = rng.integers(0, 2, count)
    eps = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), count))
    return [("poisson" if k == 0 else "aniso", int(s), float(e)) for s, k, e in zip(sides, kinds, eps)]


def build_system(H, kind, side, eps):
    return H.poisson_2d(side) if kind == "poisson" else H.anisotropic_2d(side, eps)


def bench_batched(H, ctx, ev, args):
    specs = c4_specs(args.systems)
    log(f"[bench] building {len(specs)} systems")
    systems = [build_system(H, *sp) for sp in specs]
    row_off, rp, ci, v = H.concat_systems(systems)
    N, nnz = int(row_off[-1]), len(v)
    # b = A x* per system, computed on the device in FP64
    xs = np.concatenate([H.uniform_pm1(1000 + s, len(sy[0]) - 1) for s, sy in enumerate(systems)])
    B = H.Batch(ctx, row_off, rp, ci, v)
    log(f"[bench] batch created: N={N} nnz={nnz}")
    b = np.empty(N)
    mats = []
    for s, (srp, sci, sv) in enumerate(systems):
        n = len(srp) - 1
        A = H.Csr(ctx, n, n, srp, sci, sv)
        xd = ctx.array(xs[row_off[s]:row_off[s + 1]])
        bd = ctx.array(n=n)
        A.spmv(xd, bd)
        b[row_off[s]:row_off[s + 1]] = bd.download()
        mats.append(A)
    bd = ctx.array(b)
    xd = ctx.array(n=N)
    # GPU energy counter (ROCm SMI), read around the same timed solves
    try:
        energy = H.Energy(0)
    except H.HspsError as e:
        energy = None
        log(f"[bench] no GPU energy counter: {e}")

    def avg_power(fn, min_s=3.0):
        """Device power while fn runs back to back: the energy counter ticks
        about once a second, so the average is taken between two ticks both
        observed while fn was running, at least min_s apart."""
        _, _, t_prev = energy.read()
        while True:
            fn()
            j0, _, t0 = energy.read()
            if t0 != t_prev:
                break
        while True:
            fn()
            j1, _, t1 = energy.read()
            if t1 - t0 >= min_s:
                return (j1 - j0) / (t1 - t0)

    def batched(precision, steps, warmup, batch=None):
        Bx = batch or B
        p = H.default_params(tol=1e-8, max_iter=20000, precision=precision, check_every=32,
                             batch_tile_rows=args.tile_rows)
        for _ in range(warmup):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            Bx.solve(bd, xd, p)
        times = []
        for _ in range(steps):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            ctx.sync()
            t0 = time.perf_counter()
            ok, it, cv, tr = Bx.solve(bd, xd, p)
            ctx.sync()
            times.append(time.perf_counter() - t0)
            assert ok and cv.all() and (tr <= 1e-8).all(), (precision, tr.max())
        x = xd.download()
        err = float(np.max([np.linalg.norm(x[row_off[s]:row_off[s + 1]] - xs[row_off[s]:row_off[s + 1]]) /
                            np.linalg.norm(xs[row_off[s]:row_off[s + 1]]) for s in range(len(systems))]))
        return times, it, tr, err, x

    storage = B.storage()
    sizes = np.diff(row_off).astype(np.float64)
    # algorithmic bytes of one row-iteration of the stencil-stored FP32 path:
    # K1 reads r, p_old, the presence byte and writes p_new (13 B); K2 reads
    # p, the presence byte, r, x_lo and writes r, x_lo (21 B)
    per_row = 34 if storage["stencil"] == len(systems) else None
    # Full-load iteration: the tiled systems (rows > 20,000; the smaller ones run in the persistent
    # one-workgroup kernel on a second stream) as a batch of their own, so that kernel does not
    # share the device; all of them are still iterating after 96 iterations, so (t(96) - t(32)) / 64
    # is one iteration (K1, F1, K2, F2) over all tiled rows, timed with HIP events on the stream the
    # batched kernels run on.
    stream = H.lib().hsps_ctx_stream(ctx.h)
    big = [s for s in range(len(systems)) if sizes[s] > 20000]
    t_off, t_rp, t_ci, t_v = H.concat_systems([systems[s] for s in big])
    Bt = H.Batch(ctx, t_off, t_rp, t_ci, t_v)
    bt = ctx.array(np.concatenate([b[row_off[s]:row_off[s + 1]] for s in big]))
    xt = ctx.array(n=int(t_off[-1]))

    def capped(iters):
        p = H.default_params(tol=1e-8, max_iter=iters, precision="mixed", check_every=32,
                             batch_tile_rows=args.tile_rows)
        ts = []
        for _ in range(3):
            H.check(H.lib().hsps_memset(ctx.h, xt.ptr, 0, xt.nbytes))
            ctx.sync()
            e0, e1 = ev.event(), ev.event()
            ev.record(e0, stream)
            Bt.solve(bt, xt, p)
            ev.record(e1, stream)
            ts.append(ev.elapsed_ms(e0, e1) * 1e-3)
        return float(np.median(ts))
    tiled_rows = float(t_off[-1])
    assert tiled_rows == float(sizes[sizes > 20000].sum())
    if args.full_load_only:
        # for rocprofv3 runs: full-load iterations only, plus a 256 MiB device copy that calibrates
        # the FETCH_SIZE / WRITE_SIZE counters
        nbytes = 256 << 20
        src, dst = ctx.array(n=nbytes // 8), ctx.array(n=nbytes // 8)
        H.check(H.lib().hsps_memcpy_d2d(ctx.h, dst.ptr, src.ptr, nbytes))
        ctx.sync()
        capped(32)  # warm-up (tile autotuning happens here when --tile-rows is 0)
        t_it = (capped(96) - capped(32)) / 64
        out = {"full_load_iteration_us": 1e6 * t_it, "tiled_rows": tiled_rows, "bytes_per_row_iteration": per_row,
               "full_load_achieved_GBs": per_row * tiled_rows / t_it / 1e9 if per_row else None,
               "calibration_copy_bytes": nbytes}
        for A in mats:
            A.close()
        B.close()
        Bt.close()
        return out
    log("[bench] batched mixed")
    t_mx, it_mx, tr_mx, err_mx, x_mx = batched("mixed", args.steps, args.warmup)
    log(f"[bench] batched mixed {np.mean(t_mx):.3f}s")
    row_iters = float(np.dot(sizes, it_mx.astype(np.float64)))
    t_it = (capped(96) - capped(32)) / 64
    batched_roof = {"bytes_per_row_iteration": per_row, "storage": storage, "tiled_rows": tiled_rows,
                    "full_load_iteration_us": 1e6 * t_it,
                    "full_load_achieved_GBs": per_row * tiled_rows / t_it / 1e9 if per_row else None,
                    "row_iterations": row_iters,
                    "solve_average_GBs": per_row * row_iters / float(np.mean(t_mx)) / 1e9 if per_row else None}
    if args.only_batched:
        for A in mats:
            A.close()
        B.close()
        Bt.close()
        return {"workload": "batched_only", "value": len(systems) / float(np.mean(t_mx)),
                "ms_per_step": 1e3 * float(np.mean(t_mx)), "config": {}, "fp64_baseline": None,
                "speedup_vs_fp64": None, "cpu_baseline": None, "batched_iteration": batched_roof}
    t_b64, it_b64, tr_b64, _, _ = batched("fp64", 1, 1)
    log(f"[bench] batched fp64 {np.mean(t_b64):.3f}s")

    # The same batch with every system stored as CSR, then as DIA (diagonal-major
    # values, no stencil): the general batched kernels, priced with the
    # reference's SpMV byte model (energy_monitor.hpp:166-170) in FP32 plus the
    # PCG iteration's 7 vector passes: per system and iteration
    # nnz (4 + 4) + 4 (n + 1) + 4 n + 4 n + 7 * 4 n bytes.
    storage_legs = {}
    nnz_s = np.array([len(sy[2]) for sy in systems], np.float64)
    model = nnz_s * 8 + 4 * (sizes + 1) + 8 * sizes + 28 * sizes
    for cap in ("csr", "dia"):
        os.environ["HSPS_BATCH_STORAGE"] = cap
        try:
            Bc = H.Batch(ctx, row_off, rp, ci, v)
        finally:
            del os.environ["HSPS_BATCH_STORAGE"]
        t_c, it_c, tr_c, _, _ = batched("mixed", max(1, min(args.steps, 2)), 1, Bc)
        mb = float(np.dot(model, it_c.astype(np.float64)))
        gbs = mb / float(np.mean(t_c)) / 1e9
        storage_legs[cap] = {"storage": Bc.storage(), "value": len(systems) / float(np.mean(t_c)), "unit": "solves/s",
                             "seconds": float(np.mean(t_c)), "total_iterations": int(it_c.sum()),
                             "max_true_residual": float(tr_c.max()), "model_bytes": mb, "achieved_GBs": gbs,
                             "frac_of_8TBs": gbs / HBM_PEAK_GBS}
        log(f"[bench] batched mixed, {cap} storage {np.mean(t_c):.3f}s ({gbs:.0f} GB/s by the CSR byte model)")
        Bc.close()

    # plain FP64 device PCG, one system at a time (cg_engine=0: CSR row-block SpMV and
    # BLAS-1 kernels per iteration); then the same through the one-system batched engine
    def one_at_a_time(engine):
        p64 = H.default_params(tol=1e-8, max_iter=20000, precision="fp64", check_every=32, cg_engine=engine)
        precs = [H.Precond(A, "jacobi", p64) for A in mats]
        iters = np.zeros(len(mats), int)
        t_single = []
        for rep in range(2):  # first pass warms caches/allocations, second is timed
            total = 0.0
            for s, A in enumerate(mats):
                n = A.rows
                bs = ctx.array(b[row_off[s]:row_off[s + 1]])
                xs_d = ctx.array(n=n)
                ctx.sync()
                t0 = time.perf_counter()
                ok, st = H.solve(A, bs, xs_d, "cg", precs[s], p64)
                ctx.sync()
                total += time.perf_counter() - t0
                assert ok and st.true_residual <= 1e-8
                iters[s] = st.iterations
            t_single.append(total)
        for M in precs:
            M.close()
        return t_single[-1], iters

    fp64_time, iters64 = one_at_a_time(0)
    log(f"[bench] one-at-a-time fp64 (plain kernels) {fp64_time:.3f}s")
    fp64e_time, iters64e = one_at_a_time(1)
    log(f"[bench] one-at-a-time fp64 (batched engine) {fp64e_time:.3f}s")

    # CPU oracle: the FP64 OpenMP restatement of the reference on stratified
    # samples of the same systems (every 16th / 64th by nonzeros), at all
    # physical cores and at 1 thread; its solutions also check the batched
    # mixed solutions (every result against the FP64 oracle)
    cpu, vs_oracle = None, None
    if not args.no_cpu:
        order = np.argsort([len(sy[2]) for sy in systems], kind="stable")
        sample_p, sample_1 = [int(s) for s in order[8::16]], [int(s) for s in order[32::64]]
        P, n_cpus = physical_cores()
        # protocol: 3 warm-up passes, then the median of 5 timed ones (about 25 s of CPU work in all)
        CPU_WARMUP, CPU_REPS = 3, 5
        rp_, xo = cpu_oracle_runs(specs, b, row_off, sample_p, P, CPU_WARMUP, CPU_REPS, "all physical cores",
                                  want_x=True)
        r1_, _ = cpu_oracle_runs(specs, b, row_off, sample_1, 1, CPU_WARMUP, CPU_REPS, "1 thread")
        work = np.array([len(sy[2]) for sy in systems], np.float64) * it_mx.astype(np.float64)

        def ratio(sample):  # the sample's mean work per solve over the batch's
            return float(work[sample].mean() / work.mean())
        cpu = {"value": len(sample_p) / rp_["median_s"], "unit": "solves/s", "cores": P, "kind": "port",
               "sample": f"{len(sample_p)} of the {len(systems)} systems (every 16th by nonzeros), solved by the FP64 "
                         f"OpenMP oracle at {P} threads (OMP_PROC_BIND=close, OMP_PLACES=cores; {P} physical cores of "
                         f"the {n_cpus} CPUs this process may use): {CPU_WARMUP} warm-up passes, median of {CPU_REPS} "
                         f"timed passes "
                         f"({rp_['median_s']:.2f} s); the sample's mean work per solve (nonzeros x iterations) is "
                         f"{ratio(sample_p):.3f} x the batch's, and value is the sample's own solves/s (no extrapolation)",
               "one_thread": {"value": len(sample_1) / r1_["median_s"], "unit": "solves/s", "threads": 1,
                              "systems": len(sample_1), "median_s": r1_["median_s"], "pass_seconds": r1_["pass_seconds"],
                              "work_ratio_vs_batch": ratio(sample_1),
                              "sample": f"every 64th system by nonzeros, {CPU_WARMUP} warm-ups + median of {CPU_REPS} "
                                        "passes"},
               "pass_seconds": rp_["pass_seconds"], "work_ratio_vs_batch": ratio(sample_p)}
        errs, dits = [], []
        for i, s in enumerate(sample_p):
            xo_s = xo[rp_["x_off"][i]:rp_["x_off"][i + 1]]
            errs.append(float(np.linalg.norm(x_mx[row_off[s]:row_off[s + 1]] - xo_s) / np.linalg.norm(xo_s)))
            dits.append((int(it_mx[s]), int(rp_["iterations"][i])))
        vs_oracle = {"systems": len(sample_p), "max_rel_err_vs_oracle": max(errs),
                     "max_abs_iteration_diff": max(abs(a - o) for a, o in dits),
                     "max_rel_iteration_diff": max(abs(a - o) / o for a, o in dits),
                     "iteration_ratio": float(sum(a for a, _ in dits) / sum(o for _, o in dits)),
                     "what": "batched mixed solution of each sampled system against the FP64 CPU oracle's solution "
                             "of the same b (relative 2-norm), and the batched mixed iteration counts against the "
                             "oracle's"}

    value = len(systems) / float(np.mean(t_mx))
    fp64_rate = len(systems) / fp64_time
    res = {
        "workload": f"batched_{len(systems)}_poisson_aniso_pcg_jacobi",
        "value": value,
        "ms_per_step": 1e3 * float(np.mean(t_mx)),
        "config": {"workload": f"batched {len(systems)} systems, side log-uniform 32..512, Poisson/anisotropic "
                               f"(eps log-uniform 1e-3..1), PCG+Jacobi, FP64 rel-res<=1e-8",
                   "systems": len(systems), "total_rows": N, "total_nnz": nnz, "tol": 1e-8},
        "fp64_baseline": {"value": fp64_rate, "unit": "solves/s", "mode": "plain FP64 device PCG, one system at a time",
                          "total_iterations": int(iters64.sum()), "seconds": fp64_time},
        "batched_fp64": {"value": len(systems) / float(np.mean(t_b64)), "unit": "solves/s"},
        "batched_mixed_storage": dict(storage_legs, what="the same 512 systems with the storage capped "
                                      "(HSPS_BATCH_STORAGE=csr | dia) instead of the stencil storage; achieved_GBs by "
                                      "the reference's CSR SpMV byte model (FP32) + 7 vector passes per iteration"),
        "fp64_engine_one_at_a_time": {"value": len(systems) / fp64e_time, "unit": "solves/s",
                                      "mode": "FP64 PCG+Jacobi, one system at a time, each run as a one-system "
                                              "batch (stencil storage, fused kernels)",
                                      "total_iterations": int(iters64e.sum()), "seconds": fp64e_time,
                                      "speedup_of_batched_mixed": (len(systems) / float(np.mean(t_mx))) /
                                                                  (len(systems) / fp64e_time)},
        "mixed": {"max_iterations": int(it_mx.max()), "total_iterations": int(it_mx.sum()),
                  "max_true_residual": float(tr_mx.max()), "max_rel_err_vs_xstar": err_mx,
                  "batch_time_p50_ms": 1e3 * float(np.percentile(t_mx, 50)),
                  "batch_time_p99_ms": 1e3 * float(np.percentile(t_mx, 99)),
                  "iteration_ratio_vs_fp64": float(it_mx.sum() / iters64.sum()),
                  "vs_oracle": vs_oracle},
        "speedup_decomposition": {
            "batched_mixed_over_batched_fp64": (len(systems) / float(np.mean(t_mx))) / (len(systems) / float(np.mean(t_b64))),
            "batched_fp64_over_plain_fp64_one_at_a_time": (len(systems) / float(np.mean(t_b64))) / (len(systems) / fp64_time),
            "what": "speedup_vs_fp64 = (mixed over FP64, both batched) x (batched FP64 over the plain FP64 device PCG "
                    "one system at a time)"},
        "speedup_vs_fp64": value / fp64_rate,
        "cpu_baseline": cpu,
        "batched_iteration": batched_roof,
    }
    if energy is not None and not args.no_energy:
        log("[bench] energy: device power under each workload")
        S = len(systems)
        pm = H.default_params(tol=1e-8, max_iter=20000, precision="mixed", check_every=32,
                              batch_tile_rows=args.tile_rows)
        w_mx = avg_power(lambda: B.solve(bd, xd, pm))
        p64 = H.default_params(tol=1e-8, max_iter=20000, precision="fp64", check_every=32, cg_engine=0)
        precs = [H.Precond(A, "jacobi", p64) for A in mats]
        bufs = [(ctx.array(b[row_off[s]:row_off[s + 1]]), ctx.array(n=A.rows)) for s, A in enumerate(mats)]
        nxt = [0]

        def one():
            s = nxt[0] % S
            nxt[0] += 1
            H.solve(mats[s], bufs[s][0], bufs[s][1], "cg", precs[s], p64)
            ctx.sync()
        w_64 = avg_power(one)
        for M in precs:
            M.close()
        _, _, t_prev = energy.read()
        while True:  # idle: one whole counter interval with nothing running
            j0, _, t0 = energy.read()
            if t0 != t_prev:
                break
            time.sleep(0.01)
        time.sleep(1.05)
        j1, _, t1 = energy.read()
        w_idle = (j1 - j0) / (t1 - t0) if t1 > t0 else None
        jm = w_mx * float(np.mean(t_mx)) / S
        j6 = w_64 * fp64_time / S
        res["energy"] = {"source": "GPU energy counter (ROCm SMI rsmi_dev_energy_count_get, whole device, updated "
                                   "about once a second): average power over >= 3 s of each workload run back to "
                                   "back, times its measured time per solve",
                         "batched_mixed_J_per_solve": jm, "batched_mixed_avg_W": w_mx,
                         "fp64_one_at_a_time_J_per_solve": j6, "fp64_one_at_a_time_avg_W": w_64,
                         "idle_W": w_idle,
                         "energy_ratio_fp64_over_batched": j6 / jm if jm > 0 else None,
                         "device": energy.info()}
        energy.close()
    for A in mats:
        A.close()
    B.close()
    Bt.close()
    return res


def bench_single(H, ctx, ev, args):
    side = args.grid
    n = side * side
    rp, ci, v = H.poisson_2d(side)
    A = H.Csr(ctx, n, n, rp, ci, v)
    xs = H.uniform_pm1(42, n)
    x_dev = ctx.array(xs)
    b_dev = ctx.array(n=n)
    A.spmv(x_dev, b_dev)
    b_host = b_dev.download()

    def run(precision, steps, warmup, engine=1):
        p = H.default_params(tol=1e-8, max_iter=20000, precision=precision, check_every=64, cg_engine=engine)
        M = H.Precond(A, "jacobi", p)
        xd = ctx.array(n=n)
        for _ in range(warmup):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            H.solve(A, b_dev, xd, "cg", M, p)
        times, stats = [], []
        for _ in range(steps):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            ctx.sync()
            t0 = time.perf_counter()
            ok, st = H.solve(A, b_dev, xd, "cg", M, p)
            ctx.sync()
            times.append(time.perf_counter() - t0)
            assert ok and st.true_residual <= 1e-8
            stats.append(st)
        return times, stats

    t64, s64 = run("fp64", max(1, args.steps), args.warmup, engine=0)  # plain kernel sequence
    t64e, s64e = run("fp64", max(1, args.steps), args.warmup)
    tmx, smx = run("mixed", args.steps, args.warmup)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        Ao = O.Csr.from_csr(n, n, rp, ci, v)
        t0 = time.perf_counter()
        _, sto = O.solve(Ao, b_host, "cg", "jacobi", tol=1e-8, max_iter=200)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": 1.0 / (dt / sto.iterations * s64[0].iterations), "unit": "solves/s", "cores": cores,
               "kind": "port", "sample": f"{sto.iterations} PCG iterations of the FP64 OpenMP oracle at full size "
                                         f"({dt:.1f} s), scaled to the {s64[0].iterations} a solve needs"}
    value = len(tmx) / sum(tmx)
    return {
        "workload": f"poisson2d_{side}x{side}_pcg_jacobi",
        "value": value,
        "ms_per_step": 1e3 * sum(tmx) / len(tmx),
        "config": {"workload": f"2-D Poisson {side}x{side}, PCG+Jacobi, FP64 rel-res<=1e-8", "n": n, "nnz": len(v)},
        "fp64_baseline": {"value": len(t64) / sum(t64), "unit": "solves/s", "iterations": s64[0].iterations,
                          "mode": "plain FP64 device PCG (CSR row-block SpMV + BLAS-1 kernels)"},
        "fp64_engine": {"value": len(t64e) / sum(t64e), "unit": "solves/s", "iterations": s64e[0].iterations,
                        "mode": "FP64 PCG as a one-system batch (stencil storage, fused kernels)"},
        "mixed": {"iterations": smx[0].iterations, "reliable_updates": smx[0].reliable_updates,
                  "true_residual": smx[0].true_residual, "p50_ms": 1e3 * float(np.percentile(tmx, 50)),
                  "p99_ms": 1e3 * float(np.percentile(tmx, 99))},
        "speedup_vs_fp64": value / (len(t64) / sum(t64)),
        "cpu_baseline": cpu,
    }


def bench_convdiff(H, ctx, args):
    """Convection-diffusion 1024^2 (n = 1,048,576), Pe = |beta|/eps = 100 with
    |beta| = 1 at a seeded angle, FGMRES(30) + ILU(0) to FP64 rel-res <= 1e-10:
    mixed (FP32 inner, FP64 restarts) vs FP64."""
    side = args.grid if args.grid != 2048 else 1024
    n = side * side
    th = float(np.random.default_rng(31).uniform(0.0, 2 * np.pi))
    eps = 1.0 / 100.0
    rp, ci, v = H.convection_diffusion_2d(side, eps, float(np.cos(th)), float(np.sin(th)))
    A = H.Csr(ctx, n, n, rp, ci, v)
    xd_star = ctx.array(H.uniform_pm1(77, n))
    b_dev = ctx.array(n=n)
    A.spmv(xd_star, b_dev)

    def run(precision, steps, warmup):
        # mixed: GMRES-based iterative refinement with an inner tolerance of 1e-5 per FP32 cycle
        p = H.default_params(tol=1e-10, max_iter=5000, restart_size=30, drop_tol=0.0, precision=precision,
                             fgmres_inner_rel=1e-5 if precision == "mixed" else 0.0)
        M = H.Precond(A, "ilu", p)
        xd = ctx.array(n=n)
        times, stats = [], []
        for k in range(warmup + steps):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            ctx.sync()
            t0 = time.perf_counter()
            ok, st = H.solve(A, b_dev, xd, "fgmres", M, p)
            ctx.sync()
            if k >= warmup:
                times.append(time.perf_counter() - t0)
                assert ok and st.true_residual <= 1e-10, (precision, st.true_residual)
                stats.append(st)
        M.close()
        return times, stats

    t64, s64 = run("fp64", max(1, args.steps), args.warmup)
    tmx, smx = run("mixed", args.steps, args.warmup)
    value = len(tmx) / sum(tmx)
    return {
        "workload": f"convdiff{side}_pe100_fgmres30_ilu0",
        "value": value,
        "ms_per_step": 1e3 * sum(tmx) / len(tmx),
        "config": {"workload": f"convection-diffusion {side}x{side}, Pe 100 (|beta| 1, angle {th:.3f}), "
                               "FGMRES(30) + ILU(0), FP64 rel-res<=1e-10; mixed: FP32 cycles to a 1e-5 "
                               "residual reduction, FP64 restarts", "n": n, "nnz": len(v)},
        "fp64_baseline": {"value": len(t64) / sum(t64), "unit": "solves/s", "iterations": s64[0].iterations,
                          "mode": "FP64 FGMRES(30) + ILU(0) on the device"},
        "mixed": {"iterations": smx[0].iterations, "restarts": smx[0].restarts,
                  "true_residual": smx[0].true_residual},
        "speedup_vs_fp64": value / (len(t64) / sum(t64)),
        "cpu_baseline": None,
    }


def bench_amg(H, ctx, args):
    """Poisson 2048^2 (n = 4,194,304): FGMRES(30) + SA-AMG (theta 0.25, omega 0.67
    as the reference, V(2,2)) to FP64 rel-res <= 1e-8, FP64 vs mixed; the
    hierarchy setup (host aggregation / smoothed P / Galerkin, device upload)
    is timed separately."""
    side = args.grid
    n = side * side
    rp, ci, v = H.poisson_2d(side)
    A = H.Csr(ctx, n, n, rp, ci, v)
    xs = ctx.array(H.uniform_pm1(42, n))
    b_dev = ctx.array(n=n)
    A.spmv(xs, b_dev)

    def run(precision, steps, warmup):
        p = H.default_params(tol=1e-8, max_iter=2000, restart_size=30, precision=precision)
        t0 = time.perf_counter()
        M = H.Precond(A, "amg", p)
        setup = time.perf_counter() - t0
        xd = ctx.array(n=n)
        times, stats = [], []
        for k in range(warmup + steps):
            H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
            ctx.sync()
            t0 = time.perf_counter()
            ok, st = H.solve(A, b_dev, xd, "fgmres", M, p)
            ctx.sync()
            if k >= warmup:
                times.append(time.perf_counter() - t0)
                assert ok and st.true_residual <= 1e-8, (precision, st.true_residual)
                stats.append(st)
        levels = M.amg_levels()
        prof = M.amg_setup_profile()
        M.close()
        return times, stats, (setup, prof), levels

    t64, s64, su64, lv = run("fp64", max(1, args.steps), args.warmup)
    tmx, smx, sumx, _ = run("mixed", args.steps, args.warmup)
    value = len(tmx) / sum(tmx)
    return {
        "workload": f"poisson{side}_fgmres30_amg",
        "value": value,
        "ms_per_step": 1e3 * sum(tmx) / len(tmx),
        "config": {"workload": f"2-D Poisson {side}x{side}, FGMRES(30) + SA-AMG (theta 0.25, omega 0.67, V(2,2)), "
                               "FP64 rel-res<=1e-8", "n": n, "nnz": len(v), "amg_levels": lv},
        "fp64_baseline": {"value": len(t64) / sum(t64), "unit": "solves/s", "iterations": s64[0].iterations,
                          "setup_s": su64[0]},
        "mixed": {"iterations": smx[0].iterations, "restarts": smx[0].restarts, "setup_s": sumx[0],
                  "true_residual": smx[0].true_residual},
        "amg_setup_stages_s": sumx[1],
        "speedup_vs_fp64": value / (len(t64) / sum(t64)),
        "cpu_baseline": None,
    }


def serve_specs(count, nmax, seed=77):
    """Request mix: n log-uniform in [1e3, nmax] (grid side rounded), one third
    each Poisson, anisotropic (eps log-uniform in [1e-3, 1]) and
    convection-diffusion with Pe = |beta|/eps log-uniform in [1, 1000] at a
    uniform angle (|beta| = 1, eps = 1/Pe)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        side = max(32, int(round(np.sqrt(np.exp(rng.uniform(np.log(1e3), np.log(nmax)))))))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            out.append(("poisson", side, 0.0, 0.0, 0.0))
        elif kind == 1:
            out.append(("aniso", side, float(np.exp(rng.uniform(np.log(1e-3), 0.0))), 0.0, 0.0))
        else:
            pe = float(np.exp(rng.uniform(0.0, np.log(1000.0))))
            th = float(rng.uniform(0.0, 2 * np.pi))
            out.append(("convdiff", side, 1.0 / pe, float(np.cos(th)), float(np.sin(th))))
    return out


def serve_request(H, spec, k):
    """CSR arrays and b = A x* (x* = uniform_pm1(9000 + k)) of one request."""
    import scipy.sparse as sp
    kind, side, a, bx, by = spec
    if kind == "poisson":
        rp, ci, v = H.poisson_2d(side)
    elif kind == "aniso":
        rp, ci, v = H.anisotropic_2d(side, a)
    else:
        rp, ci, v = H.convection_diffusion_2d(side, a, bx, by)
    n = side * side
    b = sp.csr_matrix((v, ci, rp), shape=(n, n)) @ H.uniform_pm1(9000 + k, n)
    return rp, ci, v, b


def bench_serve(H, ctx, args):
    """Requests arrive as a Poisson process and are submitted at their
    scheduled times; they are generated ahead of the schedule by a small
    thread pool (bounded look-ahead), so host memory holds only the requests
    in flight.  Latency is measured from the scheduled arrival to the
    server's completion timestamp."""
    from collections import deque
    from concurrent.futures import ThreadPoolExecutor
    specs = serve_specs(args.requests, args.serve_nmax)
    # mixed precision, the reference's defaults otherwise (ILUT(1e-4) at MODERATE)
    p = H.default_params(tol=1e-8, max_iter=5000, precision="mixed")
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, len(specs)))
    srv = H.Server(workers=args.workers, hbm_budget=args.hbm_budget, batch_max=64,
                   batch_rows_max=int(args.serve_batch_rows))
    # warm-up, untimed: one small request per ladder state (session, ILUT, AMG)
    warm = [("poisson", 64, 0.0, 0.0, 0.0), ("convdiff", 64, 0.01, 0.8, 0.6), ("aniso", 64, 1e-3, 0.0, 0.0)]
    for k, spec in enumerate(warm):
        srv.wait(srv.submit(*serve_request(H, spec, 10 ** 6 + k), p))
    pool = ThreadPoolExecutor(6)
    ahead, nxt = deque(), 0

    def refill():
        nonlocal nxt
        while nxt < len(specs) and len(ahead) < 32:
            ahead.append(pool.submit(serve_request, H, specs[nxt], nxt))
            nxt += 1

    refill()
    log(f"[bench] serve: {len(specs)} requests at {args.rate}/s, n in [1e3, {args.serve_nmax:.0g}], "
        f"{args.workers} workers")
    t0 = time.monotonic()  # the server's timestamps are CLOCK_MONOTONIC seconds
    tickets, rows, late = [], [], 0
    for k, ta in enumerate(arrivals):
        req = ahead.popleft().result()
        refill()
        dt = t0 + ta - time.monotonic()
        if dt > 0:
            time.sleep(dt)
        else:
            late += 1
        tickets.append(srv.submit(*req, p))
        rows.append(len(req[0]) - 1)
        del req
        if (k + 1) % 100 == 0:
            log(f"[bench] serve: {k + 1} submitted at {time.monotonic() - t0:.1f} s")
    results = []
    for t in tickets:
        ok, _, r = srv.wait(t)
        results.append((ok, r))
        if len(results) % 200 == 0:
            log(f"[bench] serve: {len(results)} done at {time.monotonic() - t0:.1f} s")
    makespan = max(r.done_s for _, r in results) - t0
    srv.close()
    pool.shutdown()
    lat = np.array([r.done_s - (t0 + ta) for (_, r), ta in zip(results, arrivals)])
    states = [int(r.state) for _, r in results]
    ok = [o for o, _ in results]

    def by_state(s):
        sel = [i for i, t in enumerate(states) if t == s]
        return {"n": len(sel), "p50_ms": 1e3 * float(np.percentile(lat[sel], 50)),
                "p99_ms": 1e3 * float(np.percentile(lat[sel], 99)),
                "mean_queue_ms": 1e3 * float(np.mean([results[i][1].queue_s for i in sel])),
                "mean_setup_ms": 1e3 * float(np.mean([results[i][1].setup_s for i in sel])),
                "mean_solve_ms": 1e3 * float(np.mean([results[i][1].solve_s for i in sel])),
                "max_rows": int(max(rows[i] for i in sel)),
                "in_session": int(sum(results[i][1].batched for i in sel))}

    return {
        "workload": f"serve_{len(specs)}_requests",
        "value": len(results) / makespan,
        "ms_per_step": 1e3 * makespan,
        "config": {"workload": f"serving {len(specs)} requests, Poisson arrivals at {args.rate}/s, n log-uniform in "
                               f"[1e3, {args.serve_nmax:.0f}], Poisson / anisotropic / convection-diffusion (Pe "
                               "1..1000), adaptive ladder, mixed precision, FP64 rel-res<=1e-8",
                   "requests": len(specs), "rate_per_s": args.rate, "workers": args.workers,
                   "session_rows_max": int(args.serve_batch_rows), "total_rows": int(sum(rows))},
        "fp64_baseline": None,
        "speedup_vs_fp64": None,
        "cpu_baseline": None,
        "serve": {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                  "mean_ms": 1e3 * float(lat.mean()), "latency_from": "scheduled arrival",
                  "converged": int(sum(ok)), "requests": len(results),
                  "max_true_residual": float(max(r.true_residual for _, r in results)),
                  "states": {H.STATE_NAMES[s]: states.count(s) for s in sorted(set(states))},
                  "in_session": int(sum(r.batched for _, r in results)),
                  "escalated": int(sum(r.escalations > 0 for _, r in results)),
                  "submitted_late": late, "makespan_s": makespan, "rows_per_s": float(sum(rows)) / makespan,
                  "by_state": {H.STATE_NAMES[s]: by_state(s) for s in sorted(set(states))}},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-energy", action="store_true", help="skip the device energy measurement")
    ap.add_argument("--workload", choices=["batched", "single", "amg", "convdiff", "serve"], default="batched")
    ap.add_argument("--requests", type=int, default=1000, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="batched workload: requests of the serving leg run after it (0: none)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="batched workload: the serving leg's arrival rate")
    ap.add_argument("--workers", type=int, default=8, help="serve: concurrent ladder solves")
    ap.add_argument("--hbm-budget", type=float, default=64e9, help="serve: admission budget (bytes)")
    ap.add_argument("--serve-nmax", type=float, default=4e6, help="serve: largest request (rows)")
    ap.add_argument("--serve-batch-rows", type=float, default=1048576,
                    help="serve: largest EASY request that joins the session (rows)")
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0, help="batched tile rows (0: autotuned on the first solve)")
    ap.add_argument("--only-batched", action="store_true",
                    help="batched mixed solves only (no FP64 or CPU baselines); for rocprofv3 runs")
    ap.add_argument("--full-load-only", action="store_true",
                    help="only the batched full-load iterations (capped solves) and a 256 MiB calibration copy; "
                         "for rocprofv3 runs")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV roofline kernel (plus a 256 MiB device copy that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters); for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    H = load_pkg()
    ev = HipEvents()
    ctx = H.Context(0)
    if args.full_load_only:
        res = bench_batched(H, ctx, ev, args)
        print(json.dumps({"batched_full_load": res}), flush=True)
        ctx.close()
        return
    log("[bench] spmv roofline")
    roof, rn, rnnz = spmv_roofline(H, ctx, ev)
    # achievable HBM streaming rate on this device (hsps_stream_probe: two FP32 reads and one write per
    # value, nontemporal 16-byte accesses, 3 x 338 MB, well above the 256 MB Infinity Cache)
    ach = C.c_double()
    H.check(H.lib().hsps_stream_probe(ctx.h, 84434944, 20, C.byref(ach)))
    achievable = {"GBs": ach.value, "kernel": "hsps_stream_probe: y = a r + b p over 84,434,944 FP32 values "
                                               "(12 B/value, nontemporal 16-byte loads and stores), median of 20"}
    log(f"[bench] achievable streaming rate {ach.value:.0f} GB/s")
    if args.spmv_only:
        nbytes = 256 << 20  # calibration copy: reads and writes exactly 256 MiB
        src, dst = ctx.array(n=nbytes // 8), ctx.array(n=nbytes // 8)
        H.check(H.lib().hsps_memcpy_d2d(ctx.h, dst.ptr, src.ptr, nbytes))
        ctx.sync()
        print(json.dumps({"spmv": roof, "n": rn, "nnz": rnnz, "calibration_copy_bytes": nbytes}), flush=True)
        ctx.close()
        return
    serve = None
    if args.workload == "batched":
        res = bench_batched(H, ctx, ev, args)
        if args.serve_requests > 0 and not args.only_batched:  # serving leg: per-request latency
            sargs = argparse.Namespace(**vars(args))
            sargs.requests = args.serve_requests
            sargs.rate = args.serve_rate
            serve = bench_serve(H, ctx, sargs)
    elif args.workload == "serve":
        res = bench_serve(H, ctx, args)
    elif args.workload == "convdiff":
        res = bench_convdiff(H, ctx, args)
    elif args.workload == "amg":
        res = bench_amg(H, ctx, args)
    else:
        res = bench_single(H, ctx, ev, args)
    r64 = roof["fp64"]
    spmv_roof = {"bound": "hbm", "kernel": "k_spmv<double,double,EpiStore> (CSR row-block SpMV, FP64), "
                                           f"Poisson 2048^2 (n={rn}, nnz={rnnz})",
                 "achieved": r64["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": r64["achieved_GBs"] / HBM_PEAK_GBS, "traffic": SPMV_FP64_HBM_BYTES,
                 "traffic_source": "rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000 per dispatch, "
                                   "calibrated on a 256 MiB copy (profiles/spmv_poisson2048_fp64_fp32.md)",
                 "algorithmic_bytes": r64["algorithmic_bytes"], "avg_launch_us": r64["avg_launch_us"],
                 "fp32_spmv": roof["fp32"],
                 "achievable": dict(achievable, frac=r64["achieved_GBs"] / achievable["GBs"])}
    bi = res.get("batched_iteration")
    if bi and bi.get("full_load_achieved_GBs"):
        alg = bi["bytes_per_row_iteration"] * bi["tiled_rows"]
        roofline = {"bound": "hbm",
                    "kernel": "batched PCG iteration at full load: k_b_k1v<float> + k_b_k2v<float> (per-system "
                              f"finishes fused in) over {bi['tiled_rows']:.0f} stencil-stored rows",
                    "achieved": bi["full_load_achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": bi["full_load_achieved_GBs"] / HBM_PEAK_GBS,
                    "traffic": BATCHED_HBM_BYTES_PER_ROW_ITER * bi["tiled_rows"],
                    "traffic_source": "rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000 of the k_b_k1v / k_b_k2v "
                                      "dispatches at full load (median per dispatch / tiled rows) "
                                      f"({BATCHED_HBM_BYTES_PER_ROW_ITER:.1f} B vs {bi['bytes_per_row_iteration']} B "
                                      "algorithmic; profiles/batched512_pcg_jacobi_kernels.md)",
                    "algorithmic_bytes": alg, "avg_launch_us": bi["full_load_iteration_us"],
                    "per": "one iteration (one launch each of K1 and K2)",
                    "achievable": dict(achievable, frac=bi["full_load_achieved_GBs"] / achievable["GBs"])}
        res["spmv_roofline"] = spmv_roof
    else:
        roofline = spmv_roof
    line = {
        "metric": "solves/s @ rel-res<=1e-8 (batched mixed vs plain FP64), p50/p99 latency",
        "value": res["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded splitmix64 x*, b = A x*)",
        "config": res["config"],
        "fp64_baseline": res["fp64_baseline"],
        "speedup_vs_fp64": res["speedup_vs_fp64"],
        "detail": {k: v for k, v in res.items() if k not in ("value", "ms_per_step", "config", "fp64_baseline",
                                                             "speedup_vs_fp64", "cpu_baseline")},
        "roofline": roofline,
        "cpu_baseline": res["cpu_baseline"],
    }
    if serve is not None:
        sv = serve["serve"]
        line["p50_latency_ms"] = sv["p50_ms"]
        line["p99_latency_ms"] = sv["p99_ms"]
        line["serve"] = dict(sv, workload=serve["config"]["workload"], requests=serve["config"]["requests"],
                             rate_per_s=serve["config"]["rate_per_s"], workers=serve["config"]["workers"],
                             total_rows=serve["config"]["total_rows"], throughput_solves_per_s=serve["value"])
    elif args.workload == "serve":
        line["p50_latency_ms"] = res["serve"]["p50_ms"]
        line["p99_latency_ms"] = res["serve"]["p99_ms"]
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
