"""GPU parity: the gfx950 path (through the C ABI) against the FP64 CPU oracle
and the reference's fixtures.

Tolerances
  SpMV FP64: |y - y_ref|_i <= 1e-13 * (|A||x|)_i   (summation order / FMA only)
  SpMV FP32: |y - y_ref|_i <= 2e-6 * (|A||x|)_i    (FP32 values and vectors)
  solves:    true FP64 rel-res <= tol; solution within 1e-6 relative of the
             oracle's; iteration counts within 5% of the reference's.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CV, SV = 0.7648421872844885, 0.644217687237691  # cos 0.7, sin 0.7


def dev_csr(H, ctx, A):
    rp, ci, v = A.arrays()
    r, c, _ = A.shape_nnz()
    return H.Csr(ctx, r, c, rp, ci, v)


def random_csr(rng, n, m, row_lens):
    rows, cols, vals = [], [], []
    for i, L in enumerate(row_lens):
        cs = rng.choice(m, size=min(L, m), replace=False)
        rows += [i] * len(cs)
        cols += list(cs)
        vals += list(rng.uniform(-1, 1, len(cs)))
    return O.Csr.from_coo(n, m, rows, cols, vals)


def absAx(A, x):
    rp, ci, v = A.arrays()
    Aabs = O.Csr.from_csr(A.shape_nnz()[0], A.shape_nnz()[1], rp, ci, np.abs(v))
    return Aabs.spmv(np.abs(x))


@pytest.mark.parametrize("case", ["poisson64", "convdiff48", "mixed_rows", "long_rows", "tiny", "empty_rows"])
def test_spmv_matches_oracle(H, ctx, case):
    rng = np.random.default_rng(11)
    if case == "poisson64":
        A = O.Csr.poisson(64)
    elif case == "convdiff48":
        A = O.Csr.convdiff(48, 0.01, CV, SV)
    elif case == "mixed_rows":  # row lengths 0..300: LDS groups of 1..64 lanes per row
        A = random_csr(rng, 3000, 5000, rng.integers(0, 300, 3000))
    elif case == "long_rows":  # rows longer than one LDS stage (2048)
        A = random_csr(rng, 40, 20000, [5000, 3, 2049, 0, 4096] + [7] * 35)
    elif case == "tiny":
        A = O.Csr.from_coo(1, 1, [0], [0], [2.0])
    else:
        A = O.Csr.from_coo(6, 6, [0, 2, 5], [1, 2, 0], [1.0, -2.0, 3.0])
    n, m, _ = A.shape_nnz()
    x = rng.uniform(-1, 1, m)
    Ad = dev_csr(H, ctx, A)
    xd, yd = ctx.array(x), ctx.array(n=n)
    Ad.spmv(xd, yd)
    y = yd.download()
    yref = A.spmv(x)
    bound = 1e-13 * absAx(A, x) + 1e-300
    assert np.all(np.abs(y - yref) <= bound), np.max(np.abs(y - yref) / bound)
    # FP32 path
    x32, y32 = ctx.array(x.astype(np.float32), dtype=np.float32), ctx.array(n=n, dtype=np.float32)
    Ad.spmv(x32, y32, precision=H.HSPS_MIXED)
    y2 = y32.download().astype(np.float64)
    assert np.all(np.abs(y2 - yref) <= 2e-6 * absAx(A, x) + 1e-30)


def test_spmv_poisson_closed_form(H, ctx):
    """A*1 for unscaled Poisson: 4 - (#neighbours) per row (0 interior)."""
    n = 256
    A = O.Csr.poisson(n)
    Ad = dev_csr(H, ctx, A)
    xd, yd = ctx.array(np.ones(n * n)), ctx.array(n=n * n)
    Ad.spmv(xd, yd)
    y = yd.download().reshape(n, n)
    expect = np.zeros((n, n))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    np.testing.assert_array_equal(y, expect)


def test_blas1_exact_small_integers(H, ctx):
    x = np.arange(1000, dtype=np.float64) % 7 - 3
    y = np.arange(1000, dtype=np.float64) % 5 - 2
    xd, yd = ctx.array(x), ctx.array(y)
    import ctypes as C
    out = C.c_double()
    H.check(H.lib().hsps_dot_f64(ctx.h, 1000, xd.ptr, yd.ptr, C.byref(out)))
    assert out.value == float(np.dot(x, y))
    H.check(H.lib().hsps_axpy_f64(ctx.h, 1000, 2.0, xd.ptr, yd.ptr))
    np.testing.assert_array_equal(yd.download(), y + 2 * x)
    H.check(H.lib().hsps_axpby_f64(ctx.h, 1000, 1.0, xd.ptr, -1.0, yd.ptr))
    np.testing.assert_array_equal(yd.download(), x - (y + 2 * x))
    H.check(H.lib().hsps_norm2_f64(ctx.h, 1000, xd.ptr, C.byref(out)))
    assert abs(out.value - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)


def _solve_dev(H, ctx, A, b, solver, pre, **kw):
    Ad = dev_csr(H, ctx, A)
    p = H.default_params(**kw)
    M = H.Precond(Ad, pre, p) if pre != "none" else None
    bd, xd = ctx.array(b), ctx.array(n=len(b))
    ok, st = H.solve(Ad, bd, xd, solver, M, p)
    return ok, st, xd.download()


def _rel(A, b, x):
    return np.linalg.norm(b - A.spmv(x)) / np.linalg.norm(b)


def test_readme_table_on_gpu(H, ctx, golden):
    """README.md:195-199 iteration counts, FP64 device path."""
    A = O.Csr.poisson(64)
    b = np.ones(4096)
    ref = golden["readme_b1_64"]
    for solver, pre, key in [("cg", "jacobi", "cg_jacobi"), ("cg", "ilu", "cg_ilu0"),
                             ("fgmres", "jacobi", "fgmres50_jacobi"), ("fgmres", "ilu", "fgmres50_ilu0"),
                             ("fgmres", "amg", "fgmres50_amg")]:
        ok, st, x = _solve_dev(H, ctx, A, b, solver, pre, tol=1e-8, max_iter=2000, restart_size=50, drop_tol=0.0)
        assert ok, key
        it = ref[key]["iters"]
        assert abs(st.iterations - it) <= max(1, 0.05 * it), (key, st.iterations, it)
        assert st.true_residual < 1e-8
        assert abs(np.linalg.norm(x) - ref[key]["x_norm"]) <= 1e-6 * ref[key]["x_norm"], key
        # SolverStats::flop_count / mem_bytes of the reference build (cg_solver.cpp:98-179,
        # fgmres_solver.cpp:196-198) whenever the iteration counts agree
        if st.iterations == it:
            assert (st.flop_count, st.mem_bytes) == (ref[key]["flop_count"], ref[key]["mem_bytes"]), key


# Mixed FGMRES cases that need one FP64 refinement cycle more than the
# reference's iterations (measured on MI355X).  An FP32 cycle leaves a
# residual gap of about u32 * || |A| |x - x0| || (FP32 Krylov vectors and
# SpMV): when its Givens estimate meets tol the FP64 true residual is still
# above tol, so a restarted cycle of two iterations follows.  With small
# counts those two exceed 5%; the test proves the gap on every listed case
# (the mixed solve stopped at the reference's count misses tol).
MIXED_REFINEMENT_CYCLE = {("poisson16_s1", "fgmres30_ilu0"): 20, ("poisson32_s7", "fgmres30_ilu0"): 31,
                          ("convdiff32_s5", "fgmres30_jacobi"): 18, ("convdiff64_s9", "fgmres30_jacobi"): 22}


@pytest.mark.parametrize("idx", range(6))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_seeded_solves(H, ctx, golden, idx, precision):
    c = golden["seeded"][idx]
    A = O.Csr.poisson(c["n"]) if c["kind"] == "poisson" else O.Csr.convdiff(c["n"], c["eps"], c["bx"], c["by"])
    xs = O.uniform(c["seed"], A.rows)
    b = A.spmv(xs)
    kw = dict(tol=1e-8, max_iter=20000, restart_size=30, drop_tol=0.0)
    bad = []
    for key, solver, pre in [("cg_jacobi", "cg", "jacobi"), ("fgmres30_jacobi", "fgmres", "jacobi"),
                             ("fgmres30_ilu0", "fgmres", "ilu"), ("fgmres30_amg", "fgmres", "amg")]:
        if key not in c:
            continue
        ref = c[key]
        ok, st, x = _solve_dev(H, ctx, A, b, solver, pre, precision=precision, **kw)
        assert ok, (key, precision)
        assert _rel(A, b, x) < 1e-8
        xo, sto = O.solve(A, b, solver, pre, **kw)
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo), (key, precision)
        it = ref["iters"]
        print(f"{c['case']} {key} {precision}: {st.iterations} iterations (reference {it})", flush=True)
        # within 5% of the reference (at least 1), FP64 and mixed alike
        # (mixed: total inner iterations)
        if abs(st.iterations - it) > max(1, 0.05 * it):
            exc = MIXED_REFINEMENT_CYCLE.get((c["case"], key)) if precision == "mixed" else None
            if exc is None or st.iterations != exc:
                bad.append((key, precision, st.iterations, it))
            else:  # the gap: stopped at the reference's count, the mixed solve misses tol
                _, st2, x2 = _solve_dev(H, ctx, A, b, solver, pre, precision=precision,
                                        **dict(kw, max_iter=it))
                assert st2.true_residual > 1e-8 and _rel(A, b, x2) > 1e-8, (key, st2.true_residual)
        if precision == "fp64" and st.iterations == it:  # the reference's flop model, counted alike
            assert (st.flop_count, st.mem_bytes) == (ref["flop_count"], ref["mem_bytes"]), key
    assert not bad, bad


def test_amg_hierarchy_bit_exact(H, ctx, golden):
    for c in golden["amg"]:
        name = c["case"]
        if not name.startswith("poisson") or name == "poisson180":
            continue
        n = int(name[7:].split("_")[0])
        A = O.Csr.poisson(n)
        Ad = dev_csr(H, ctx, A)
        M = H.Precond(Ad, "amg", H.default_params(amg_strong=c["theta"]))
        assert M.amg_levels() == len(c["levels"]), name
        oa = O.Amg(A, amg_strong=c["theta"])
        for lv, L in enumerate(c["levels"]):
            assert M.amg_level_info(lv) == (L["rows"], L["nnz"]), (name, lv)
            if lv + 1 < len(c["levels"]):
                np.testing.assert_array_equal(M.amg_aggregates(lv), oa.aggregates(lv))


def _arrow(side, stride=3, c=0.05):
    """Poisson side^2 plus a hub row/column coupled to every stride-th node: a
    P row and a Galerkin row reaching thousands of coarse columns, and an R
    row (the hub's aggregate) with thousands of entries."""
    rp, ci, v = O.Csr.poisson(side).arrays()
    n = side * side
    rows = np.repeat(np.arange(n), np.diff(rp))
    link = np.arange(0, n, stride)
    r = np.concatenate([rows, np.full(len(link), n), link, [n]])
    cc = np.concatenate([ci, link, np.full(len(link), n), [n]])
    vv = np.concatenate([v, np.full(len(link), -c), np.full(len(link), -c), [1.0 + c * len(link)]])
    return O.Csr.from_coo(n + 1, n + 1, r, cc, vv)


def _hub_row(side, stride, diag=1e16):
    """Poisson side^2 plus one row (no column) coupled to every stride-th node
    with a huge diagonal: its smoothed-P entries fall under the 1e-14 drop, but
    its row of A P and of the Galerkin product reach thousands of columns of a
    coarse space wider than one LDS window."""
    rp, ci, v = O.Csr.poisson(side).arrays()
    n = side * side
    rows = np.repeat(np.arange(n), np.diff(rp))
    link = np.arange(0, n, stride)
    r = np.concatenate([rows, np.full(len(link), n), [n]])
    cc = np.concatenate([ci, link, [n]])
    vv = np.concatenate([v, np.full(len(link), -1.0), [diag]])
    return O.Csr.from_coo(n + 1, n + 1, r, cc, vv)


def _graph_laplacian(rng, n, max_deg):
    """Random weighted graph Laplacian + I: rows of 2..max_deg entries."""
    r, c, w = [], [], []
    for i in range(n):
        for j in rng.choice(n, size=int(rng.integers(1, max_deg // 2)), replace=False):
            if j != i:
                r.append(i); c.append(int(j)); w.append(float(rng.uniform(0.1, 1.0)))
    r, c, w = np.array(r), np.array(c), np.array(w)
    rr = np.concatenate([r, c])
    cc = np.concatenate([c, r])
    ww = np.concatenate([-w, -w])
    deg = np.bincount(rr, weights=-ww, minlength=n)
    rr = np.concatenate([rr, np.arange(n)])
    cc = np.concatenate([cc, np.arange(n)])
    ww = np.concatenate([ww, deg + 1.0])
    return O.Csr.from_coo(n, n, rr, cc, ww)


@pytest.mark.parametrize("case", ["poisson100", "poisson64_t05", "aniso64", "convdiff64", "graph4000", "arrow60", "hub130"])
def test_amg_device_hierarchy_bitwise(H, ctx, case):
    """Every level's A, P and R built on the device equal the oracle's
    restatement of the reference's setup bit for bit (aggregates included)."""
    theta = 0.5 if case.endswith("_t05") else 0.25
    if case.startswith("poisson"):
        A = O.Csr.poisson(int(case[7:].split("_")[0]))
    elif case == "aniso64":
        A = O.Csr.aniso(64, 0.01)
    elif case == "convdiff64":
        A = O.Csr.convdiff(64, 0.01, CV, SV)
    elif case == "graph4000":
        A = _graph_laplacian(np.random.default_rng(3), 4000, 160)
    elif case == "arrow60":
        A = _arrow(60, stride=4)
    else:
        A = _hub_row(130, 13)
    M = H.Precond(dev_csr(H, ctx, A), "amg", H.default_params(amg_strong=theta))
    oa = O.Amg(A, amg_strong=theta)
    L = M.amg_levels()
    assert L == oa.levels(), (case, L, oa.levels())
    for lv in range(L):
        whats = (["A"] if lv else []) + (["P", "R"] if lv + 1 < L else [])
        for w in whats:
            rp, ci, v = M.amg_level_matrix(lv, w)
            orp, oci, ov = oa.matrix(lv, w).arrays()
            msg = f"{case} level {lv} {w}"
            np.testing.assert_array_equal(rp, orp, err_msg=msg)
            np.testing.assert_array_equal(ci, oci, err_msg=msg)
            np.testing.assert_array_equal(v.view(np.int64), np.ascontiguousarray(ov, np.float64).view(np.int64),
                                          err_msg=msg)
        if lv + 1 < L:
            np.testing.assert_array_equal(M.amg_aggregates(lv), oa.aggregates(lv), err_msg=f"{case} agg {lv}")
    prof = M.amg_setup_profile()
    assert all(t >= 0 for t in prof.values()) and prof["galerkin"] > 0


def test_amg_vcycle_matches_oracle(H, ctx):
    A = O.Csr.poisson(48)
    Ad = dev_csr(H, ctx, A)
    M = H.Precond(Ad, "amg")
    r = O.uniform(5, A.rows)
    rd, zd = ctx.array(r), ctx.array(n=A.rows)
    M.apply(rd, zd)
    z = zd.download()
    zo = O.Amg(A).apply(r)
    assert np.linalg.norm(z - zo) <= 1e-10 * np.linalg.norm(zo)


def test_ilu_apply_matches_oracle(H, ctx):
    for A, tau in [(O.Csr.convdiff(24, 0.05, 1.0, 0.5), 0.0), (O.Csr.convdiff(16, 0.1, 5.0, 0.0), 1e-4)]:
        Ad = dev_csr(H, ctx, A)
        M = H.Precond(Ad, "ilu", H.default_params(drop_tol=tau))
        r = O.uniform(3, A.rows)
        rd, zd = ctx.array(r), ctx.array(n=A.rows)
        M.apply(rd, zd)
        zo = O.precond_apply(A, r, "ilu", drop_tol=tau)
        assert np.linalg.norm(zd.download() - zo) <= 1e-12 * np.linalg.norm(zo)


@pytest.mark.parametrize("n", [23, 25, 1024, 1030, 2048, 2049])
def test_ilu_apply_level_shapes(H, ctx, n):
    """ILU(0) of an n x n convection-diffusion grid in natural order: the
    sweeps run 2n - 1 anti-diagonal levels up to n rows wide.  n = 23 / 24
    (above) / 25 leave 5 / 7 / 1 levels after the last whole round of the
    FP64 sweep's 8 prefetched levels; n = 1024 fills the 1024-lane sweep at
    its widest level; n = 1030 and 2048 take two rows per lane (2048 fills
    them); n = 2049 exceeds that (the sweep that loops over a level's rows)."""
    A = O.Csr.convdiff(n, 0.05, 1.0, 0.5)
    Ad = dev_csr(H, ctx, A)
    M = H.Precond(Ad, "ilu", H.default_params(drop_tol=0.0))
    r = O.uniform(11, A.rows)
    rd, zd = ctx.array(r), ctx.array(n=A.rows)
    M.apply(rd, zd)
    z = zd.download()
    zo = O.precond_apply(A, r, "ilu", drop_tol=0.0)
    assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)
    M.apply(rd, zd)  # a second apply reuses the level-ordered buffers
    assert np.array_equal(zd.download(), z)


def test_reference_unit_known_answers(H, ctx):
    """tests/test_solvers.cpp:42-76, 151-167 of the reference."""
    A = O.Csr.from_coo(1, 1, [0], [0], [2.0])
    ok, st, x = _solve_dev(H, ctx, A, np.array([4.0]), "cg", "jacobi", tol=1e-9, max_iter=3000)
    assert ok and abs(x[0] - 2.0) < 1e-10
    D = O.Csr.from_coo(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2.0, 3.0, 5.0, 7.0])
    ok, st, x = _solve_dev(H, ctx, D, np.array([4.0, 9.0, 15.0, 21.0]), "cg", "jacobi", tol=1e-9)
    np.testing.assert_allclose(x, [2, 3, 3, 3], atol=1e-10)
    D3 = O.Csr.from_coo(3, 3, [0, 1, 2], [0, 1, 2], [3.0, 6.0, 9.0])
    ok, st, x = _solve_dev(H, ctx, D3, np.array([3.0, 6.0, 9.0]), "fgmres", "jacobi", tol=1e-9, restart_size=40)
    assert ok
    np.testing.assert_allclose(x, [1, 1, 1], atol=1e-10)
    # zero RHS: converged, 0 iterations (test_solvers.cpp:117-129)
    P = O.Csr.poisson(8)
    ok, st, x = _solve_dev(H, ctx, P, np.zeros(64), "cg", "jacobi")
    assert ok and st.iterations == 0


def test_errors_as_status_codes(H, ctx):
    A = O.Csr.from_coo(3, 3, [0, 1, 2], [0, 1, 2], [1.0, 0.0, 2.0])
    Ad = dev_csr(H, ctx, A)
    with pytest.raises(H.HspsError) as e:
        H.Precond(Ad, "jacobi")
    assert e.value.code == -3 and "zero diagonal at row 1" in str(e.value)
    with pytest.raises(H.HspsError) as e:
        H.Precond(Ad, "ilu", H.default_params(drop_tol=0.0))
    assert e.value.code == -4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_engine_matches_plain_kernels(H, ctx, precision):
    """PCG + Jacobi on one system runs as a one-system batch (diagonal /
    stencil storage, fused kernels) unless it is a large CSR-stored matrix.
    Both paths meet the oracle's bars; the CSR fallback is the plain path
    itself (bit-identical)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    n = 30000
    i = np.repeat(np.arange(n), 3)
    j = rng.integers(0, n, 3 * n)
    G = sp.coo_matrix((-np.ones(3 * n), (i, j)), shape=(n, n)).tocsr()
    G = G + G.T
    G.setdiag(0)
    G.eliminate_zeros()
    L = (G + sp.diags(-np.asarray(G.sum(axis=1)).ravel() + 0.05)).tocsr()
    L.sort_indices()
    # the last case has a 1100-row halo: the engine cuts one staging chunk per tile (2200 rows) and
    # runs the five-point kernels with 512-thread workgroups (a 17.6 KB FP32 / 35 KB FP64 window)
    cases = [O.Csr.poisson(200), O.Csr.aniso(150, 0.05), O.Csr.poisson(40),
             O.Csr.from_csr(n, n, L.indptr.astype(np.int32), L.indices.astype(np.int32), L.data),
             O.Csr.poisson(1100)]
    for k, A in enumerate(cases):
        b = A.spmv(O.uniform(40 + k, A.rows))
        xo, so = O.solve(A, b, "cg", "jacobi", tol=1e-8, max_iter=20000)
        res = {}
        for engine in (0, 1):
            ok, st, x = _solve_dev(H, ctx, A, b, "cg", "jacobi", tol=1e-8, max_iter=20000, precision=precision,
                                   cg_engine=engine)
            assert ok and _rel(A, b, x) < 1e-8 and abs(st.true_residual - _rel(A, b, x)) <= 1e-3 * st.true_residual
            assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo), (k, engine)
            assert abs(st.iterations - so.iterations) <= max(1, 0.05 * so.iterations), (k, engine, precision,
                                                                                        st.iterations, so.iterations)
            res[engine] = (st.iterations, x)
        if k == 3:  # large and CSR-stored: the plain kernels run either way
            assert res[0][0] == res[1][0]
            np.testing.assert_array_equal(res[0][1], res[1][1])


def test_mixed_ilu_two_rows_per_lane(H, ctx):
    """Conv-diff 1100^2: ILU levels up to 1100 rows wide, so the FP32 sweeps
    of the mixed solve take two rows per lane.  The mixed solve reaches the
    FP64 tolerance in the FP64 device solve's iterations (within 10% or 2),
    and both match the FP64 oracle's solution."""
    A = O.Csr.convdiff(1100, 0.05, 1.0, 0.5)
    b = A.spmv(O.uniform(21, A.rows))
    xo, so = O.solve(A, b, "fgmres", "ilu", tol=1e-10, max_iter=2000, restart_size=30, drop_tol=0.0)
    its = {}
    for precision in ("fp64", "mixed"):
        ok, st, x = _solve_dev(H, ctx, A, b, "fgmres", "ilu", tol=1e-10, max_iter=2000, restart_size=30,
                               drop_tol=0.0, precision=precision, fgmres_inner_rel=1e-5 if precision == "mixed" else 0.0)
        assert ok and _rel(A, b, x) < 1e-10, precision
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo), precision
        its[precision] = st.iterations
    assert abs(its["fp64"] - so.iterations) <= max(1, 0.05 * so.iterations), (its, so.iterations)
    assert abs(its["mixed"] - so.iterations) <= max(1, 0.05 * so.iterations), (its, so.iterations)


def test_mixed_fgmres_inner_tolerance(H, ctx):
    """GMRES-based refinement to 1e-10: an FP32 cycle that stops at a 1e-5
    residual reduction (hsps_params.fgmres_inner_rel) and lets the FP64
    restart refine needs no more iterations than cycles run to tol, and stays
    close to the FP64 oracle's count."""
    A = O.Csr.convdiff(256, 0.05, 0.6, 0.8)  # the FP64 oracle needs 26 iterations
    b = A.spmv(O.uniform(9, A.rows))
    xo, so = O.solve(A, b, "fgmres", "ilu", tol=1e-10, max_iter=2000, restart_size=30, drop_tol=0.0)
    runs = {}
    for inner in (0.0, 1e-5):
        ok, st, x = _solve_dev(H, ctx, A, b, "fgmres", "ilu", tol=1e-10, max_iter=2000, restart_size=30,
                               drop_tol=0.0, precision="mixed", fgmres_inner_rel=inner)
        assert ok and _rel(A, b, x) < 1e-10 and st.true_residual < 1e-10, inner
        runs[inner] = st.iterations
    assert runs[1e-5] <= runs[0.0], runs
    assert abs(runs[1e-5] - so.iterations) <= max(1, 0.05 * so.iterations), (runs, so.iterations)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jacobi_from_other_values_keeps_its_diagonal(H, ctx, precision):
    """A Jacobi set up before hsps_csr_update_values keeps the D^-1 it was
    built with, as the reference's CG applies M's stored inverse diagonal
    (cg_solver.cpp:79, 147, jacobi.cpp:40-51): the one-system engine (which
    rebuilds D^-1 from A) must not take such a solve, so cg_engine = 1 and 0
    give the same bits; a Jacobi set up afresh takes the engine again."""
    A = O.Csr.poisson(64)
    rp, ci, v = A.arrays()
    n = A.rows
    rows = np.repeat(np.arange(n), np.diff(rp))
    v2 = v + np.where(rows == ci, O.uniform(5, len(v)) + 1.5, 0.0)  # a new diagonal in [4.5, 6.5]
    A2 = O.Csr.from_csr(n, n, rp, ci, v2)
    b = A2.spmv(O.uniform(6, n))
    Ad = dev_csr(H, ctx, A)
    M_old = H.Precond(Ad, "jacobi")
    Ad.update_values(v2)
    res = {}
    for engine in (0, 1):
        p = H.default_params(tol=1e-8, max_iter=5000, precision=precision, cg_engine=engine)
        bd, xd = ctx.array(b), ctx.array(n=n)
        ok, st = H.solve(Ad, bd, xd, "cg", M_old, p)
        assert ok and _rel(A2, b, xd.download()) < 1e-8
        res[engine] = (st.iterations, xd.download())
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])
    M_new = H.Precond(Ad, "jacobi")
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision)
    bd, xd = ctx.array(b), ctx.array(n=n)
    ok, st = H.solve(Ad, bd, xd, "cg", M_new, p)
    xo, so = O.solve(A2, b, "cg", "jacobi", tol=1e-8, max_iter=5000)
    assert ok and abs(st.iterations - so.iterations) <= max(1, 0.05 * so.iterations)
    assert np.linalg.norm(xd.download() - xo) <= 1e-6 * np.linalg.norm(xo)
    M_old.close()
    M_new.close()


def test_preconditioner_keeps_its_matrix_alive(H, ctx):
    """hsps_precond_create retains the matrix handle: destroying the matrix
    first leaves the preconditioner usable (the AMG V-cycle reads level 0 at
    apply time); after hsps_csr_update_values an AMG apply is refused (its
    coarse levels were built from the old values)."""
    A = O.Csr.poisson(48)
    Ad = dev_csr(H, ctx, A)
    M = H.Precond(Ad, "amg")
    r = O.uniform(5, A.rows)
    rd, zd = ctx.array(r), ctx.array(n=A.rows)
    M.apply(rd, zd)
    z0 = zd.download()
    Ad.close()  # the caller's handle goes; the preconditioner's reference remains
    zd2 = ctx.array(n=A.rows)
    M.apply(rd, zd2)
    np.testing.assert_array_equal(zd2.download(), z0)
    M.close()  # the last reference: the matrix is freed here
    B = dev_csr(H, ctx, A)
    MB = H.Precond(B, "amg")
    B.update_values(A.arrays()[2] * 2.0)
    with pytest.raises(H.HspsError) as e:
        MB.apply(rd, zd)
    assert e.value.code == -1 and "values changed" in str(e.value)
    MB.close()
    B.close()


def _dense_csr(D):
    r, c = np.nonzero(D)
    return O.Csr.from_coo(D.shape[0], D.shape[1], r, c, D[r, c])


@pytest.mark.parametrize("case", ["poisson5", "one", "rand7", "rand64", "rand65", "pivot130", "rand300", "singular40",
                                  "rand1300", "rand2100"])
def test_amg_coarse_solve_matches_gepp(H, ctx, case):
    """The coarsest level's dense solve, factored on the device (blocked LU
    with partial pivoting, FP64 MFMA updates), against the oracle's restatement
    of direct_solve (amg.cpp:260-291).  With amg_levels = 1 (or <= 50 rows) the
    hierarchy is the matrix alone and the preconditioner is that solve.
    Sizes around the 64-row padding and the 16-column panels; pivot130 needs a
    row swap in every column; singular40 has a zero row and column (the pivot
    below 1e-14 is skipped and that unknown set to 0, as the reference does);
    rand2100 starts with panels taller than the register panel's 2048 rows
    (the global-memory panel)."""
    rng = np.random.default_rng(7)
    if case == "poisson5":
        A = O.Csr.poisson(5)
    elif case == "one":
        A = O.Csr.from_coo(1, 1, [0], [0], [3.0])
    elif case.startswith("rand"):
        n = int(case[4:])
        D = rng.uniform(-1, 1, (n, n)) * (rng.uniform(size=(n, n)) < (0.3 if n < 2000 else 0.02))
        D += np.diag(rng.uniform(2, 4, n) * np.sign(rng.uniform(-1, 1, n)) + 0.0)
        A = _dense_csr(D)
    elif case == "pivot130":
        n = 130
        D = np.diag(rng.uniform(0.01, 0.02, n)) + np.diag(rng.uniform(3, 4, n - 1), -1)
        D[0, -1] = 3.5
        D += 0.05 * rng.uniform(-1, 1, (n, n)) * (rng.uniform(size=(n, n)) < 0.1)
        A = _dense_csr(D)
    else:
        n = 40
        D = rng.uniform(-1, 1, (n, n)) * (rng.uniform(size=(n, n)) < 0.3) + np.diag(rng.uniform(3, 4, n))
        D[17, :] = 0.0
        D[:, 17] = 0.0
        A = _dense_csr(D)
    n = A.rows
    Ad = dev_csr(H, ctx, A)
    M = H.Precond(Ad, "amg", H.default_params(amg_levels=1))
    assert M.amg_levels() == 1
    for seed in (1, 2):
        r = O.uniform(seed, n)
        rd, zd = ctx.array(r), ctx.array(n=n)
        M.apply(rd, zd)
        z = zd.download()
        zo = O.gepp_solve(A, r)
        assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo), (case, np.linalg.norm(z - zo) / np.linalg.norm(zo))
        if case == "singular40":
            assert z[17] == 0.0 and zo[17] == 0.0
    prof = M.amg_setup_profile()
    print(f"{case}: coarse operator of {n} rows built in {1e3 * prof['coarse']:.2f} ms", flush=True)
    M.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_user_preconditioner_through_the_c_abi(H, ctx, precision):
    """hsps_precond_create_user: a caller-written apply (here host numpy code
    between device copies: z = D^-1 r) inside the device CG and FGMRES loops
    (cg_solver.cpp:79, 147; fgmres_solver.cpp:112).  FP64: the built-in
    Jacobi's iterations (the plain kernels) and its solution to 1e-12; mixed
    (FP32 vectors converted around the FP64 callback): the built-in Jacobi's
    mixed iterations within 5%; both: the oracle's solution to 1e-6."""
    import ctypes as C
    calls = [0]
    for solver, A in (("cg", O.Csr.aniso(40, 0.1)), ("fgmres", O.Csr.convdiff(40, 0.05, 0.8, 0.6))):
        n = A.rows
        rp, ci, v = A.arrays()
        rows = np.repeat(np.arange(n), np.diff(rp))
        dinv = 1.0 / v[rows == ci]
        b = A.spmv(O.uniform(12, n))
        Ad = dev_csr(H, ctx, A)

        def apply(d_r, d_z, stream, n=n, dinv=dinv):
            r = np.empty(n)
            H.check(H.lib().hsps_memcpy_d2h(ctx.h, r.ctypes.data_as(C.c_void_p), C.c_void_p(d_r), 8 * n))
            z = dinv * r
            H.check(H.lib().hsps_memcpy_h2d(ctx.h, C.c_void_p(d_z), z.ctypes.data_as(C.c_void_p), 8 * n))
            calls[0] += 1
        U = H.UserPrecond(ctx, n, apply)
        p = H.default_params(tol=1e-8, max_iter=5000, restart_size=30, precision=precision, cg_engine=0)
        J = H.Precond(Ad, "jacobi", p)
        res = {}
        for name, M in (("jacobi", J), ("user", U)):
            bd, xd = ctx.array(b), ctx.array(n=n)
            ok, st = H.solve(Ad, bd, xd, solver, M, p)
            assert ok and st.true_residual < 1e-8, (solver, name)
            res[name] = (st.iterations, xd.download())
        xo, so = O.solve(A, b, solver, "jacobi", tol=1e-8, max_iter=5000, restart_size=30)
        ju, jj = res["user"][0], res["jacobi"][0]
        print(f"{solver} {precision}: user {ju}, built-in Jacobi {jj}, oracle {so.iterations} iterations", flush=True)
        if precision == "fp64":  # the same iterations, the same solution to rounding
            assert ju == jj, (solver, ju, jj)
            assert np.linalg.norm(res["user"][1] - res["jacobi"][1]) <= 1e-12 * np.linalg.norm(res["jacobi"][1])
            assert abs(ju - so.iterations) <= max(1, 0.05 * so.iterations), (solver, ju, so.iterations)
        else:  # FP32 z rounded from the FP64 callback instead of an FP32 product
            assert abs(ju - jj) <= max(1, 0.05 * jj), (solver, ju, jj)
        assert np.linalg.norm(res["user"][1] - xo) <= 1e-6 * np.linalg.norm(xo), solver
        J.close()
        if solver == "fgmres":
            def failing(d_r, d_z, stream):
                raise RuntimeError("no")
            F = H.UserPrecond(ctx, n, failing)
            with pytest.raises(H.HspsError) as e:
                H.solve(Ad, ctx.array(b), ctx.array(n=n), "fgmres", F, H.default_params(tol=1e-8, precision=precision))
            assert e.value.code == -1 and "user preconditioner" in str(e.value)
            F.close()
        U.close()
    assert calls[0] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fgmres_ilut_matches_oracle(H, ctx, precision):
    """FGMRES(30) + ILUT(1e-4), the reference factory's MODERATE state
    (solver_factory.cpp:73-79, ilu.cpp:94-220), on conv-diff 128^2: the factors
    come from the sparse-scan ILUT (bit-identical to the dense scan, checked on
    the CPU), the sweeps run their fill-in dependencies."""
    A = O.Csr.convdiff(128, 0.01, CV, SV)
    b = A.spmv(O.uniform(19, A.rows))
    kw = dict(tol=1e-8, max_iter=2000, restart_size=30, drop_tol=1e-4)
    ok, st, x = _solve_dev(H, ctx, A, b, "fgmres", "ilu", precision=precision, **kw)
    xo, so = O.solve(A, b, "fgmres", "ilu", **kw)
    print(f"ILUT {precision}: {st.iterations} iterations (oracle {so.iterations})", flush=True)
    assert ok and _rel(A, b, x) < 1e-8
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert abs(st.iterations - so.iterations) <= max(1, 0.05 * so.iterations), (st.iterations, so.iterations)
