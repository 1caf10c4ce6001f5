"""Write tests/golden/fullsize_oracle.json: the FP64 CPU oracle's results on
the full-size single-system workloads, which take the oracle minutes (too
long for a GPU test to recompute):

* Poisson 2048^2 (n = 4,194,304), x* = splitmix64 U[-1,1] (seed 42),
  b = A x* (the oracle's sequential SpMV): PCG + Jacobi to 1e-8 and
  FGMRES(30) + SA-AMG (theta 0.25, omega 0.67, V(2,2)) to 1e-8;
* convection-diffusion 1024^2 at Pe = |beta| / eps = 100 (|beta| = 1,
  eps = 0.01) with the seeded angle theta = U[0, 2 pi) of numpy's
  default_rng(31), x* seed 77: FGMRES(30) + ILU(0) to 1e-10.

For each: iterations, converged, the recursive / Givens residual, the flop
model's counters, ||x|| and every 1024th entry of x (the GPU tests compare
their solution there).  Plus the SA-AMG hierarchy of Poisson 2048^2
(theta 0.25): per level rows, nnz, the aggregate map and A / P / R as
digests (fixture_digest.py); the oracle needs about two minutes for it.
Entries already in the file are kept (delete the file to redo them).
Test infrastructure only; run on the CPU.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
from fixture_digest import csr_digest, digest  # noqa: E402

STRIDE = 1024


def c3_angle():
    return float(np.random.default_rng(31).uniform(0.0, 2 * np.pi))


def record(x, st, t):
    return {"iters": int(st.iterations), "converged": int(st.converged), "restarts": int(st.restarts),
            "final_residual": float(st.final_residual), "flop_count": int(st.flop_count),
            "mem_bytes": int(st.mem_bytes), "x_norm": float(np.linalg.norm(x)),
            "x_stride": STRIDE, "x_samples": [float(v) for v in x[::STRIDE]], "oracle_seconds": round(t, 1)}


def amg_hierarchy(A, theta=0.25):
    amg = O.Amg(A, amg_strong=theta)
    L = amg.levels()
    levels = []
    for lv in range(L):
        Al = amg.matrix(lv, "A")
        rows, _, nnz = Al.shape_nnz()
        ent = {"rows": rows, "nnz": nnz}
        if lv:
            ent["A"] = csr_digest(*Al.arrays())
        if lv + 1 < L:
            agg = amg.aggregates(lv)
            ent["agg"] = digest(agg)
            ent["n_agg"] = int(agg.max()) + 1
            ent["P"] = csr_digest(*amg.matrix(lv, "P").arrays())
            ent["R"] = csr_digest(*amg.matrix(lv, "R").arrays())
        levels.append(ent)
    return {"theta": theta, "levels": levels}


def main():
    path = os.path.join(HERE, "fullsize_oracle.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    out.setdefault("threads", int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    side = 2048
    A = O.Csr.poisson(side)
    b = A.spmv(O.uniform(42, side * side))
    for key, solver, pre in [("poisson2048_cg_jacobi", "cg", "jacobi"), ("poisson2048_fgmres30_amg", "fgmres", "amg")]:
        if key in out:
            continue
        t0 = time.perf_counter()
        x, st = O.solve(A, b, solver, pre, tol=1e-8, max_iter=20000, restart_size=30)
        out[key] = record(x, st, time.perf_counter() - t0)
        print(key, st.iterations, f"{time.perf_counter() - t0:.1f} s", flush=True)
    if "poisson2048_amg_hierarchy" not in out:
        t0 = time.perf_counter()
        out["poisson2048_amg_hierarchy"] = amg_hierarchy(A)
        print("hierarchy", f"{time.perf_counter() - t0:.1f} s", flush=True)
    key = "convdiff1024_pe100_fgmres30_ilu0"
    if key not in out:
        th = c3_angle()
        C = O.Csr.convdiff(1024, 0.01, float(np.cos(th)), float(np.sin(th)))
        b = C.spmv(O.uniform(77, 1024 * 1024))
        t0 = time.perf_counter()
        x, st = O.solve(C, b, "fgmres", "ilu", tol=1e-10, max_iter=5000, restart_size=30, drop_tol=0.0)
        out[key] = dict(record(x, st, time.perf_counter() - t0), theta=th)
        print("convdiff", st.iterations, flush=True)
    with open(path, "w") as f:
        json.dump(out, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
