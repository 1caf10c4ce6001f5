"""Batched PCG+Jacobi (hsps_batch_*) against the FP64 CPU oracle, per system.

Bars: every system's true FP64 rel-res <= tol; its solution within 1e-6
relative of the oracle's single-system PCG+Jacobi solution; iteration counts
within 5% (FP64) / 10% (mixed, reliable-update refinement) of the oracle's;
and a system's result is bit-identical alone and inside a mixed batch (same
tile size, and for a tiled five-point system the same first row modulo 4),
since tiles never straddle systems and each system's reductions run over its
own tiles in a fixed order.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def make_systems(H, specs):
    systems, oracles = [], []
    for kind, n, eps in specs:
        if kind == "poisson":
            arr, A = H.poisson_2d(n), O.Csr.poisson(n)
        elif kind == "aniso":
            arr, A = H.anisotropic_2d(n, eps), O.Csr.aniso(n, eps)
        else:
            arr, A = H.convection_diffusion_2d(n, 1.0, 0.0, 0.0), O.Csr.convdiff(n, 1.0, 0.0, 0.0)
        systems.append(arr)
        oracles.append(A)
    return systems, oracles


SPECS = [("poisson", 32, 0), ("aniso", 40, 0.01), ("poisson", 1, 0), ("aniso", 17, 1.0), ("poisson", 64, 0),
         ("aniso", 64, 0.001), ("poisson", 5, 0), ("sym", 24, 0), ("poisson", 100, 0), ("aniso", 33, 0.2)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle(H, ctx, precision):
    systems, oracles = make_systems(H, SPECS)
    B = H.Batch.from_systems(ctx, systems)
    rng_b = []
    for s, A in enumerate(oracles):
        xs = O.uniform(100 + s, A.rows)
        rng_b.append(A.spmv(xs))
    b = np.concatenate(rng_b)
    bd, xd = ctx.array(b), ctx.array(n=len(b))
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision)
    ok, iters, conv, tr = B.solve(bd, xd, p)
    x = xd.download()
    assert ok and conv.all(), (iters, conv, tr)
    off = B.row_off
    for s, A in enumerate(oracles):
        xs_dev = x[off[s]:off[s + 1]]
        bs = rng_b[s]
        rel = np.linalg.norm(bs - A.spmv(xs_dev)) / np.linalg.norm(bs)
        assert rel < 1e-8 and abs(rel - tr[s]) <= 1e-3 * rel + 1e-16, (s, rel, tr[s])
        xo, st = O.solve(A, bs, "cg", "jacobi", tol=1e-8, max_iter=5000)
        assert np.linalg.norm(xs_dev - xo) <= 1e-6 * np.linalg.norm(xo), s
        assert abs(iters[s] - st.iterations) <= max(1, 0.05 * st.iterations), (s, precision, iters[s], st.iterations)


def test_batch_alone_equals_in_mix(H, ctx):
    systems, oracles = make_systems(H, SPECS)
    b_all = [A.spmv(O.uniform(7 + s, A.rows)) for s, A in enumerate(oracles)]
    p = H.default_params(tol=1e-8, max_iter=5000, precision="mixed", batch_tile_rows=128)
    B = H.Batch.from_systems(ctx, systems)
    bd, xd = ctx.array(np.concatenate(b_all)), ctx.array(n=B.n)
    ok, it_mix, _, _ = B.solve(bd, xd, p)
    x_mix = xd.download()
    assert ok
    for s in [0, 4, 5, 8]:
        Bs = H.Batch.from_systems(ctx, [systems[s]])
        bs, xs = ctx.array(b_all[s]), ctx.array(n=len(b_all[s]))
        ok1, it1, _, _ = Bs.solve(bs, xs, p)
        assert ok1 and it1[0] == it_mix[s]
        np.testing.assert_array_equal(xs.download(), x_mix[B.row_off[s]:B.row_off[s + 1]])


def test_batch_zero_rhs_and_warm_start(H, ctx):
    systems, oracles = make_systems(H, [("poisson", 16, 0), ("poisson", 20, 0)])
    B = H.Batch.from_systems(ctx, systems)
    b = np.concatenate([np.zeros(256), oracles[1].spmv(O.uniform(3, 400))])
    bd, xd = ctx.array(b), ctx.array(n=len(b))
    ok, it, cv, tr = B.solve(bd, xd, H.default_params(tol=1e-8))
    assert ok and it[0] == 0 and cv.all()
    # warm start from the converged x: 0 iterations
    ok, it2, cv2, _ = B.solve(bd, xd, H.default_params(tol=1e-8))
    assert ok and it2[1] == 0


def test_batch_rejects_bad_input(H, ctx):
    rp, ci, v = H.poisson_2d(4)
    ci_bad = ci.copy()
    ci_bad[0] = 20  # outside the 16-row block of system 0
    with pytest.raises(H.HspsError) as e:
        H.Batch(ctx, np.array([0, 16, 32], np.int32), np.concatenate([rp[:-1], rp + rp[-1]]).astype(np.int32),
                np.concatenate([ci_bad, ci + 16]), np.concatenate([v, v]))
    assert e.value.code == -1


def varcoef_2d(H, n, seed):
    """5-point Poisson with a row-varying diagonal 4 + u_i, u_i in [0, 1):
    nonzeros on 5 diagonals, values not constant along them (DIA storage)."""
    rp, ci, v = H.poisson_2d(n)
    v = v.copy()
    u = np.abs(O.uniform(seed, n * n))
    rows = np.repeat(np.arange(n * n), np.diff(rp))
    v[ci == rows] += u
    return rp, ci, v


def graph_spd(n, seed, deg=3):
    """Laplacian of a pseudo-random graph plus 0.05 I: columns scattered over
    the whole system (CSR storage)."""
    rng = np.random.default_rng(seed)
    a = np.repeat(np.arange(n), deg)
    b = rng.integers(0, n, n * deg)
    keep = a != b
    a, b = a[keep], b[keep]
    i = np.concatenate([a, b])
    j = np.concatenate([b, a])
    import scipy.sparse as sp
    L = sp.coo_matrix((-np.ones(len(i)), (i, j)), shape=(n, n)).tocsr()
    L.sum_duplicates()
    d = -np.asarray(L.sum(axis=1)).ravel() + 0.05
    L = (L + sp.diags(d)).tocsr()
    L.sort_indices()
    return L.indptr.astype(np.int32), L.indices.astype(np.int32), L.data.astype(np.float64)


def csr_of(M):
    import scipy.sparse as sp
    M = sp.csr_matrix(M)
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)


def poisson_3d(n):
    """7-point Poisson on an n^3 grid: seven diagonals, offsets up to n^2."""
    import scipy.sparse as sp
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity(n)
    return csr_of(sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T))


def wide_stencil(n, w):
    """Five diagonals at offsets 0, +-1, +-w (w wider than any LDS staging
    window): 4 on the diagonal, -1 off."""
    import scipy.sparse as sp
    return csr_of(sp.diags([-np.ones(n - w), -np.ones(n - 1), 4 * np.ones(n), -np.ones(n - 1), -np.ones(n - w)],
                           [-w, -1, 0, 1, w]))


def grid_2d(W, Hg, eps):
    """Anisotropic five-point operator on a grid of Hg rows of W points:
    -eps to the west and east, -1 to the north and south, 2 + 2 eps on the
    diagonal, no coupling across a grid-row end (stencil storage; 2-D tiles
    when W > 512)."""
    import scipy.sparse as sp
    Tx = sp.diags([-eps * np.ones(W - 1), 2 * eps * np.ones(W), -eps * np.ones(W - 1)], [-1, 0, 1])
    Ty = sp.diags([-np.ones(Hg - 1), 2 * np.ones(Hg), -np.ones(Hg - 1)], [-1, 0, 1])
    return csr_of(sp.kron(sp.identity(Hg), Tx) + sp.kron(Ty, sp.identity(W)))


def tiled_systems(H):
    """Systems above the persistent-kernel size (tiled path) and below it,
    one of each storage kind, plus stencils with seven diagonals and with a
    halo wider than the kernels' LDS staging window."""
    arrs = [H.poisson_2d(160), varcoef_2d(H, 150, 5), graph_spd(24000, 1), H.anisotropic_2d(170, 0.02),
            H.poisson_2d(30), graph_spd(500, 2), varcoef_2d(H, 40, 6), poisson_3d(30), wide_stencil(30000, 2000)]
    kinds = ["stencil", "dia", "csr", "stencil", "stencil", "csr", "dia", "stencil", "stencil"]
    oracles = [O.Csr.from_csr(len(rp) - 1, len(rp) - 1, rp, ci, v) for rp, ci, v in arrs]
    return arrs, kinds, oracles


@pytest.mark.parametrize("storage", ["auto", "dia", "csr"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_tiled_storage_kinds(H, ctx, monkeypatch, storage, precision):
    """Tiled path on every storage kind, each capped by HSPS_BATCH_STORAGE,
    against the oracle's single-system PCG+Jacobi."""
    arrs, kinds, oracles = tiled_systems(H)
    if storage != "auto":
        monkeypatch.setenv("HSPS_BATCH_STORAGE", storage)
    B = H.Batch.from_systems(ctx, arrs)
    cap = {"auto": 2, "dia": 1, "csr": 0}[storage]
    rank = {"csr": 0, "dia": 1, "stencil": 2}
    want = {"csr": 0, "dia": 0, "stencil": 0}
    for k in kinds:
        want[["csr", "dia", "stencil"][min(rank[k], cap)]] += 1
    assert B.storage() == want
    bs = [A.spmv(O.uniform(300 + s, A.rows)) for s, A in enumerate(oracles)]
    bd, xd = ctx.array(np.concatenate(bs)), ctx.array(n=B.n)
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision)
    ok, iters, conv, tr = B.solve(bd, xd, p)
    x = xd.download()
    assert ok and conv.all(), (iters, conv, tr)
    for s, A in enumerate(oracles):
        xs_dev = x[B.row_off[s]:B.row_off[s + 1]]
        rel = np.linalg.norm(bs[s] - A.spmv(xs_dev)) / np.linalg.norm(bs[s])
        assert rel < 1e-8 and abs(rel - tr[s]) <= 1e-3 * rel + 1e-16, (s, rel, tr[s])
        xo, st = O.solve(A, bs[s], "cg", "jacobi", tol=1e-8, max_iter=5000)
        assert np.linalg.norm(xs_dev - xo) <= 1e-6 * np.linalg.norm(xo), s
        assert abs(iters[s] - st.iterations) <= max(1, 0.05 * st.iterations), (s, precision, iters[s], st.iterations)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_tiled_alone_equals_in_mix(H, ctx, precision):
    """Tiles never straddle systems and every per-system reduction runs over
    the system's own tiles in a fixed order: a tiled system's iterates are
    bit-identical alone and inside a batch."""
    arrs, _, oracles = tiled_systems(H)
    b_all = [A.spmv(O.uniform(11 + s, A.rows)) for s, A in enumerate(oracles)]
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision, batch_tile_rows=512)
    B = H.Batch.from_systems(ctx, arrs)
    bd, xd = ctx.array(np.concatenate(b_all)), ctx.array(n=B.n)
    ok, it_mix, _, _ = B.solve(bd, xd, p)
    x_mix = xd.download()
    assert ok
    for s in [0, 1, 2, 8]:  # stencil, DIA, CSR, and the halo that takes the 512-thread five-point kernels
        Bs = H.Batch.from_systems(ctx, [arrs[s]])
        bs, xs = ctx.array(b_all[s]), ctx.array(n=len(b_all[s]))
        ok1, it1, _, _ = Bs.solve(bs, xs, p)
        assert ok1 and it1[0] == it_mix[s]
        np.testing.assert_array_equal(xs.download(), x_mix[B.row_off[s]:B.row_off[s + 1]])


def test_batch_solves_with_new_vectors_each_call(H, ctx):
    """The iteration chunks are replayed as hipGraphs whose kernels hold the
    b and x pointers: a later solve on other vectors must not replay them."""
    arrs, _, oracles = tiled_systems(H)
    arrs, oracles = arrs[:2], oracles[:2]
    B = H.Batch.from_systems(ctx, arrs)
    p = H.default_params(tol=1e-8, max_iter=5000, precision="mixed", batch_tile_rows=1024)
    outs = []
    for seed in (1, 2, 3):
        bs = [A.spmv(O.uniform(seed * 10 + s, A.rows)) for s, A in enumerate(oracles)]
        bd, xd = ctx.array(np.concatenate(bs)), ctx.array(n=B.n)
        ok, it, cv, tr = B.solve(bd, xd, p)
        x = xd.download()
        assert ok and (tr <= 1e-8).all()
        for s, A in enumerate(oracles):
            xs = x[B.row_off[s]:B.row_off[s + 1]]
            assert np.linalg.norm(bs[s] - A.spmv(xs)) <= 1e-8 * np.linalg.norm(bs[s])
        outs.append((bd, xd))  # keep the buffers alive: each call gets new addresses


def test_batch_params_change_between_calls(H, ctx):
    """Replayed iteration chunks carry the kernel arguments they were captured
    with: a later call with another iteration cap, tolerance or precision on
    the same vectors must run with its own (regression: the graph cache was
    keyed on wave / precision / chunk only)."""
    arrs, _, oracles = tiled_systems(H)
    arrs, oracles = [arrs[0], arrs[3], arrs[4]], [oracles[0], oracles[3], oracles[4]]
    B = H.Batch.from_systems(ctx, arrs)
    bs = [A.spmv(O.uniform(40 + s, A.rows)) for s, A in enumerate(oracles)]
    bd, xd = ctx.array(np.concatenate(bs)), ctx.array(n=B.n)

    def run(**kw):
        H.check(H.lib().hsps_memset(ctx.h, xd.ptr, 0, xd.nbytes))
        return B.solve(bd, xd, H.default_params(batch_tile_rows=1024, **kw))

    ok, it8, cv, tr = run(tol=1e-8, max_iter=5000, precision="mixed")
    assert ok and (tr <= 1e-8).all()
    ok, it10, cv, tr = run(tol=1e-8, max_iter=10, precision="mixed")
    assert not cv.any() and (it10 <= 10).all(), it10
    ok, it11, cv, tr = run(tol=1e-11, max_iter=5000, precision="mixed")
    assert ok and (tr <= 1e-11).all() and (it11 > it8).all(), (it8, it11)
    ok, it8b, cv, tr = run(tol=1e-8, max_iter=5000, precision="mixed")
    assert ok and np.array_equal(it8b, it8)


def dominant_stencil(n, w, d=8.0):
    """Five-point pattern (offsets +-1, +-w) with diagonal d > 4: CG+Jacobi
    converges in a few dozen iterations."""
    import scipy.sparse as sp
    return csr_of(sp.diags([-np.ones(n - w), -np.ones(n - 1), d * np.ones(n), -np.ones(n - 1), -np.ones(n - w)],
                           [-w, -1, 0, 1, w]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_grid_shrinks_to_live_systems(H, ctx, precision):
    """Systems that halt leave the launch grids: once the live systems' tiles
    are at most half of those launched, the host rebuilds the active tile
    records and system list.  Three fast systems (354 of the 404 tiles) halt
    within two chunks and the slow one runs on alone for hundreds of
    iterations; every system gives the same iterations and bits as alone."""
    arrs = [dominant_stencil(60000, 250), dominant_stencil(60000, 240), dominant_stencil(60000, 260),
            H.poisson_2d(160)]
    oracles = [O.Csr.from_csr(len(rp) - 1, len(rp) - 1, rp, ci, v) for rp, ci, v in arrs]
    b_all = [A.spmv(O.uniform(70 + s, A.rows)) for s, A in enumerate(oracles)]
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision, batch_tile_rows=512)
    B = H.Batch.from_systems(ctx, arrs)
    assert B.storage() == {"csr": 0, "dia": 0, "stencil": 4}
    bd, xd = ctx.array(np.concatenate(b_all)), ctx.array(n=B.n)
    ok, it_mix, conv, tr = B.solve(bd, xd, p)
    x_mix = xd.download()
    assert ok and conv.all(), (it_mix, conv, tr)
    assert max(it_mix[:3]) < 64 and it_mix[3] > 4 * 32, it_mix  # fast ones halt within two chunks of 32
    for s, A in enumerate(oracles):
        xs = x_mix[B.row_off[s]:B.row_off[s + 1]]
        rel = np.linalg.norm(b_all[s] - A.spmv(xs)) / np.linalg.norm(b_all[s])
        assert rel < 1e-8, (s, rel)
        xo, st = O.solve(A, b_all[s], "cg", "jacobi", tol=1e-8, max_iter=5000)
        assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo), s
        assert abs(it_mix[s] - st.iterations) <= max(1, 0.05 * st.iterations), (s, precision, it_mix[s],
                                                                                 st.iterations)
        Bs = H.Batch.from_systems(ctx, [arrs[s]])
        bs, x1 = ctx.array(b_all[s]), ctx.array(n=len(b_all[s]))
        ok1, it1, _, _ = Bs.solve(bs, x1, p)
        assert ok1 and it1[0] == it_mix[s], (s, it1[0], it_mix[s])
        np.testing.assert_array_equal(x1.download(), xs)


def grid_systems(H):
    """Five-point systems on full grids 516, 1,028 and 600 points wide (the
    wide-halo five-point kernels), grid heights that are not multiples of 8,
    and small systems between them so that they start at rows 1, 2 and 3
    modulo 4 (16-byte groups that straddle a system's first row); then a
    narrow grid and a wide stencil that is not a grid."""
    arrs = [H.poisson_2d(31), grid_2d(516, 41, 0.3), H.poisson_2d(5), grid_2d(1028, 21, 1.0), H.poisson_2d(3),
            grid_2d(600, 50, 0.05), H.poisson_2d(150), wide_stencil(30000, 2000)]
    oracles = [O.Csr.from_csr(len(rp) - 1, len(rp) - 1, rp, ci, v) for rp, ci, v in arrs]
    bs = [A.spmv(O.uniform(700 + s, A.rows)) for s, A in enumerate(oracles)]
    return arrs, oracles, bs


GRIDS = [1, 3, 5]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_wide_grids_at_unaligned_rows(H, ctx, precision):
    """Wide full grids starting at rows 1, 2, 3 mod 4 against the oracle; the
    same system alone (first row 0) converges alike.  (Its bits may differ:
    the five-point kernels deal rows to lanes by 16-byte groups, so a
    system's dot products keep their order at a fixed start modulo 4.)"""
    arrs, oracles, bs = grid_systems(H)
    p = H.default_params(tol=1e-8, max_iter=5000, precision=precision, batch_tile_rows=512)
    B = H.Batch.from_systems(ctx, arrs)
    assert B.storage()["stencil"] == len(arrs)
    assert [B.row_off[k] % 4 for k in GRIDS] == [1, 2, 3]
    bd, xd = ctx.array(np.concatenate(bs)), ctx.array(n=B.n)
    ok, iters, conv, tr = B.solve(bd, xd, p)
    x = xd.download()
    assert ok and conv.all(), (iters, conv, tr)
    for s, A in enumerate(oracles):
        xs_dev = x[B.row_off[s]:B.row_off[s + 1]]
        rel = np.linalg.norm(bs[s] - A.spmv(xs_dev)) / np.linalg.norm(bs[s])
        assert rel < 1e-8 and abs(rel - tr[s]) <= 1e-3 * rel + 1e-16, (s, rel, tr[s])
        xo, st = O.solve(A, bs[s], "cg", "jacobi", tol=1e-8, max_iter=5000)
        assert np.linalg.norm(xs_dev - xo) <= 1e-6 * np.linalg.norm(xo), s
        assert abs(iters[s] - st.iterations) <= max(1, 0.05 * st.iterations), (s, precision, iters[s], st.iterations)
    for s in GRIDS:
        Bs = H.Batch.from_systems(ctx, [arrs[s]])
        b1, x1 = ctx.array(bs[s]), ctx.array(n=len(bs[s]))
        ok1, it1, _, _ = Bs.solve(b1, x1, p)
        assert ok1 and abs(int(it1[0]) - int(iters[s])) <= 1, (s, it1[0], iters[s])
        xa = x[B.row_off[s]:B.row_off[s + 1]]
        assert np.linalg.norm(x1.download() - xa) <= 1e-6 * np.linalg.norm(xa), s
