"""GPU parity at the workloads' full sizes, against the FP64 CPU oracle's
results in tests/golden/fullsize_oracle.json (written on the CPU by
tests/golden/gen_fullsize_oracle.py: the oracle needs 40 s to 2 min per case,
too long to rerun inside a GPU test).

* Poisson 2048^2 (n = 4,194,304), b = A x*, x* = splitmix64 U[-1,1] seed 42:
  PCG + Jacobi (FP64 plain kernels, FP64 and mixed through the one-system
  engine) and FGMRES(30) + SA-AMG (FP64 and mixed) to 1e-8;
* the SA-AMG hierarchy of Poisson 2048^2: per level rows, nnz, aggregate map
  and A / P / R bit-identical to the oracle's (by digest);
* convection-diffusion 1024^2 at Pe = |beta| / eps = 100 (|beta| = 1,
  eps = 0.01, seeded angle), x* seed 77: FGMRES(30) + ILU(0) to 1e-10, FP64
  and mixed (FP32 cycles to a 1e-5 reduction, FP64 restarts).

Bars (the north star's): FP64 true rel-res <= tol; the solution within 1e-6
relative of the oracle's (on every 1024th entry and in norm); iteration
counts within 5% (at least 1) of the oracle's; in FP64 the flop model's
counters equal the oracle's (= the reference's) whenever the iteration
counts are equal.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from fixture_digest import csr_digest, digest  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_oracle.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def poisson2048(H, ctx):
    side = 2048
    n = side * side
    rp, ci, v = H.poisson_2d(side)
    A = O.Csr.from_csr(n, n, rp, ci, v)
    b = A.spmv(O.uniform(42, n))
    Ad = H.Csr(ctx, n, n, rp, ci, v)
    yield A, Ad, b
    Ad.close()


def _check(ref, st, x, b, A, tol, label, flops=True, xstar=None):
    """xstar (the exact solution) switches the solution bar from "within
    1e-6 of the oracle's" to "at least as close to x* as the oracle's": for a
    solve whose iterates leave the oracle's path on a system whose solution
    the tolerance fixes only to about cond(A) * tol."""
    print(f"{label}: iterations {st.iterations} (oracle {ref['iters']}), true rel-res {st.true_residual:.3e}",
          flush=True)
    rel = np.linalg.norm(b - A.spmv(x)) / np.linalg.norm(b)
    assert rel <= tol and st.true_residual <= tol, (label, rel)
    xs = x[::ref["x_stride"]]
    xo = np.array(ref["x_samples"])
    d = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
    if xstar is None:
        assert d <= 1e-6, (label, d)
        assert abs(np.linalg.norm(x) - ref["x_norm"]) <= 1e-6 * ref["x_norm"], label
    else:
        e, eo = (np.linalg.norm(v - xstar[::ref["x_stride"]]) / np.linalg.norm(xstar[::ref["x_stride"]])
                 for v in (xs, xo))
        print(f"{label}: |x - x_oracle| {d:.2e}, |x - x*| {e:.2e}, |x_oracle - x*| {eo:.2e} (relative)", flush=True)
        assert e <= eo, (label, e, eo)
    assert abs(st.iterations - ref["iters"]) <= max(1, 0.05 * ref["iters"]), (label, st.iterations, ref["iters"])
    if flops and st.iterations == ref["iters"]:
        assert (st.flop_count, st.mem_bytes) == (ref["flop_count"], ref["mem_bytes"]), label


@pytest.mark.parametrize("precision,engine", [("fp64", 0), ("fp64", 1), ("mixed", 1)])
def test_poisson2048_pcg_jacobi(H, ctx, full, poisson2048, precision, engine):
    A, Ad, b = poisson2048
    p = H.default_params(tol=1e-8, max_iter=20000, precision=precision, cg_engine=engine)
    M = H.Precond(Ad, "jacobi", p)
    bd, xd = ctx.array(b), ctx.array(n=len(b))
    ok, st = H.solve(Ad, bd, xd, "cg", M, p)
    assert ok
    _check(full["poisson2048_cg_jacobi"], st, xd.download(), b, A, 1e-8, f"C2 PCG {precision} engine {engine}",
           flops=precision == "fp64")
    M.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_poisson2048_fgmres_amg(H, ctx, full, poisson2048, precision):
    A, Ad, b = poisson2048
    p = H.default_params(tol=1e-8, max_iter=2000, restart_size=30, precision=precision)
    M = H.Precond(Ad, "amg", p)
    bd, xd = ctx.array(b), ctx.array(n=len(b))
    ok, st = H.solve(Ad, bd, xd, "fgmres", M, p)
    assert ok
    # Mixed: the V-cycle in FP32 changes the (flexible) Krylov space and one
    # FP64 refinement cycle follows; Poisson 2048^2 has cond(A) ~ 1.7e6, so the
    # solutions at rel-res 1e-8 differ by ~4e-5 (measured) while the oracle's own
    # error against x* is 2.3e-4: the bar is "at least as close to x* as the oracle".
    _check(full["poisson2048_fgmres30_amg"], st, xd.download(), b, A, 1e-8, f"C2 FGMRES+AMG {precision}",
           flops=precision == "fp64", xstar=O.uniform(42, len(b)) if precision == "mixed" else None)
    M.close()


def test_poisson2048_amg_hierarchy(H, ctx, full, poisson2048):
    """Every level of the device-built hierarchy at 2048^2 (10 levels, 130 M
    nonzeros in A, P and R) equals the oracle's bit for bit."""
    _, Ad, _ = poisson2048
    ref = full["poisson2048_amg_hierarchy"]
    M = H.Precond(Ad, "amg", H.default_params(amg_strong=ref["theta"]))
    L = M.amg_levels()
    assert L == len(ref["levels"])
    for lv, R in enumerate(ref["levels"]):
        assert M.amg_level_info(lv) == (R["rows"], R["nnz"]), lv
        if lv + 1 < L:
            agg = M.amg_aggregates(lv)
            assert int(agg.max()) + 1 == R["n_agg"] and digest(agg) == R["agg"], lv
        for w in (["A"] if lv else []) + (["P", "R"] if lv + 1 < L else []):
            got = csr_digest(*M.amg_level_matrix(lv, w))
            assert got == R[w], (lv, w)
        print(f"C2 AMG level {lv}: rows {R['rows']} nnz {R['nnz']} equal", flush=True)
    M.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_convdiff1024_pe100_fgmres_ilu0(H, ctx, full, precision):
    ref = full["convdiff1024_pe100_fgmres30_ilu0"]
    side = 1024
    n = side * side
    th = ref["theta"]
    assert th == float(np.random.default_rng(31).uniform(0.0, 2 * np.pi))  # the benchmark's seeded angle
    rp, ci, v = H.convection_diffusion_2d(side, 0.01, float(np.cos(th)), float(np.sin(th)))
    A = O.Csr.from_csr(n, n, rp, ci, v)
    b = A.spmv(O.uniform(77, n))
    Ad = H.Csr(ctx, n, n, rp, ci, v)
    p = H.default_params(tol=1e-10, max_iter=5000, restart_size=30, drop_tol=0.0, precision=precision,
                         fgmres_inner_rel=1e-5 if precision == "mixed" else 0.0)
    M = H.Precond(Ad, "ilu", p)
    bd, xd = ctx.array(b), ctx.array(n=n)
    ok, st = H.solve(Ad, bd, xd, "fgmres", M, p)
    assert ok
    _check(ref, st, xd.download(), b, A, 1e-10, f"C3 FGMRES+ILU(0) {precision}", flops=precision == "fp64")
    M.close()
    Ad.close()


def test_poisson2048_coarsest_solve_matches_gepp(H, ctx, poisson2048):
    """The 861-row coarsest operator of the 2048^2 hierarchy (padded to 896:
    14 panels), factored on the device, against the oracle's direct_solve
    restatement; and the time of that factorisation."""
    _, Ad, _ = poisson2048
    M = H.Precond(Ad, "amg", H.default_params())
    L = M.amg_levels()
    rp, ci, v = M.amg_level_matrix(L - 1, "A")
    n = len(rp) - 1
    t_coarse = M.amg_setup_profile()["coarse"]
    M.close()
    Ac = O.Csr.from_csr(n, n, rp, ci, v)
    Cd = H.Csr(ctx, n, n, rp, ci, v)
    Mc = H.Precond(Cd, "amg", H.default_params(amg_levels=1))
    r = O.uniform(3, n)
    rd, zd = ctx.array(r), ctx.array(n=n)
    Mc.apply(rd, zd)
    zo = O.gepp_solve(Ac, r)
    err = np.linalg.norm(zd.download() - zo) / np.linalg.norm(zo)
    print(f"C2 coarsest level: {n} rows, device factorisation {1e3 * t_coarse:.2f} ms, rel. diff to GEPP {err:.2e}",
          flush=True)
    assert err <= 1e-12
    Mc.close()
    Cd.close()
