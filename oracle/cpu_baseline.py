"""CPU baseline child of bench.py: the FP64 OpenMP oracle (a restatement of the
reference's PCG + Jacobi, cg_solver.cpp:37-191) timed on a stratified sample
of the batched workload.

TEST / BASELINE INFRASTRUCTURE ONLY: run by bench.py's cpu_baseline leg as a
child process (so that OMP_NUM_THREADS / OMP_PROC_BIND / OMP_PLACES take
effect in a fresh OpenMP runtime); never part of the product path.

    python oracle/cpu_baseline.py IN.npz OUT.json [--solutions OUT_X.npz]

IN.npz holds the sample: kinds (0 Poisson, 1 anisotropic), sides, eps, the
concatenated right-hand sides `b` with offsets `b_off`, and `warmup` /
`reps`.  Each pass solves every sample system once (tol 1e-8, max 20000
iterations, as the GPU workload); the JSON gives every timed pass's seconds,
their median, and per-system iterations; --solutions stores the solutions of
the last pass.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402


def main():
    inp, out = sys.argv[1], sys.argv[2]
    xs_path = sys.argv[sys.argv.index("--solutions") + 1] if "--solutions" in sys.argv else None
    d = np.load(inp)
    kinds, sides, eps = d["kinds"], d["sides"], d["eps"]
    b, off = d["b"], d["b_off"]
    warmup, reps = int(d["warmup"]), int(d["reps"])
    mats = [O.Csr.poisson(int(s)) if k == 0 else O.Csr.aniso(int(s), float(e)) for k, s, e in zip(kinds, sides, eps)]

    def one_pass():
        its, xs = [], []
        t0 = time.perf_counter()
        for i, A in enumerate(mats):
            x, st = O.solve(A, b[off[i]:off[i + 1]], "cg", "jacobi", tol=1e-8, max_iter=20000)
            assert st.converged, i
            its.append(int(st.iterations))
            xs.append(x)
        return time.perf_counter() - t0, its, xs

    for _ in range(warmup):
        one_pass()
    times = []
    for _ in range(reps):
        t, its, xs = one_pass()
        times.append(t)
    res = {"threads": int(os.environ.get("OMP_NUM_THREADS", "0")), "proc_bind": os.environ.get("OMP_PROC_BIND"),
           "places": os.environ.get("OMP_PLACES"), "warmup": warmup, "pass_seconds": times,
           "median_s": float(np.median(times)), "systems": len(mats), "iterations": its}
    with open(out, "w") as f:
        json.dump(res, f)
    if xs_path:
        np.savez(xs_path, x=np.concatenate(xs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
