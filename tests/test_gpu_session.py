"""Sessions (hsps_session_*): systems join a running batched PCG + Jacobi at
chunk boundaries and leave it when they end.

Bars: a system's iterations and solution are bit-identical whether it runs
alone, joins a session mid-solve beside others, or reuses a slot and ranges
another system freed; and as hsps_batch_solve of that system alone at the
same tile rows (tiled path).  Against the FP64 oracle's single-system
PCG + Jacobi: true FP64 rel-res <= tol, solution within 1e-6 relative,
iterations within 5%, as the batch tests.
"""
import numpy as np
import pytest

import oracle as O
from test_gpu_batch import graph_spd, grid_2d, varcoef_2d, wide_stencil

pytestmark = pytest.mark.gpu


def systems(H):
    arrs = [H.poisson_2d(160), graph_spd(24000, 1), varcoef_2d(H, 150, 5), H.poisson_2d(30), graph_spd(500, 2),
            wide_stencil(30000, 2000), H.anisotropic_2d(120, 0.02), grid_2d(600, 50, 0.05)]
    oracles = [O.Csr.from_csr(len(rp) - 1, len(rp) - 1, rp, ci, v) for rp, ci, v in arrs]
    bs = [A.spmv(O.uniform(500 + s, A.rows)) for s, A in enumerate(oracles)]
    return arrs, oracles, bs


def run_to_end(S, pending, results, sizes, max_steps=5000):
    for _ in range(max_steps):
        if not pending:
            return
        for slot in S.step():
            k = pending.pop(slot)
            results[k] = S.result(slot, sizes[k])
    raise AssertionError("session did not drain")


def params(H, precision, graphs=1):
    return H.default_params(tol=1e-8, max_iter=5000, precision=precision, batch_tile_rows=512,
                            batch_graphs=graphs)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_session_join_mid_solve_equals_alone(H, ctx, precision):
    arrs, oracles, bs = systems(H)
    sizes = [A.rows for A in oracles]
    p = params(H, precision)
    S = H.Session(ctx, 8, sum(sizes), sum(len(a[1]) for a in arrs), p)
    res, pending = {}, {}
    # system 0 starts; the others join one or a few chunks later, mid-solve
    pending[S.add(*arrs[0], bs[0])] = 0
    order = [1, 2, 3, 4, 5, 6, 7]
    step = 0
    while pending or order:
        if order and step in (2, 3, 5, 6, 9, 12, 14):
            k = order.pop(0)
            pending[S.add(*arrs[k], bs[k])] = k
        for slot in S.step():
            k = pending.pop(slot)
            res[k] = S.result(slot, sizes[k])
        step += 1
        assert step < 5000
    assert S.live() == 0
    # each alone in the same (now empty) session: slots and ranges reused
    for k in range(len(arrs)):
        alone = {}
        run_to_end(S, {S.add(*arrs[k], bs[k]): k}, alone, sizes)
        x1, it1, cv1, tr1 = alone[k]
        x0, it0, cv0, tr0 = res[k]
        assert cv0 and cv1 and it0 == it1, (k, it0, it1)
        np.testing.assert_array_equal(x0, x1)
        assert tr0 == tr1
    S.close()
    for k, A in enumerate(oracles):
        x, it, cv, tr = res[k]
        rel = np.linalg.norm(bs[k] - A.spmv(x)) / np.linalg.norm(bs[k])
        assert rel < 1e-8 and abs(rel - tr) <= 1e-3 * rel + 1e-16, (k, rel, tr)
        xo, st = O.solve(A, bs[k], "cg", "jacobi", tol=1e-8, max_iter=5000)
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo), k
        assert abs(it - st.iterations) <= max(1, 0.05 * st.iterations), (k, precision, it, st.iterations)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_session_equals_batch_solve(H, ctx, precision):
    """A tiled system (above the persistent-kernel size) in a session gives the
    bits of hsps_batch_solve of that system alone at the same tile rows."""
    arrs, oracles, bs = systems(H)
    p = params(H, precision)
    for k in [0, 1, 2, 5, 7]:
        n = oracles[k].rows
        S = H.Session(ctx, 2, n, len(arrs[k][1]), p)
        out = {}
        run_to_end(S, {S.add(*arrs[k], bs[k]): k}, out, {k: n})
        S.close()
        B = H.Batch.from_systems(ctx, [arrs[k]])
        bd, xd = ctx.array(bs[k]), ctx.array(n=n)
        ok, it, _, tr = B.solve(bd, xd, p)
        assert ok and it[0] == out[k][1], (k, it[0], out[k][1])
        np.testing.assert_array_equal(xd.download(), out[k][0])


def test_session_graphs_match_direct_launches(H, ctx):
    arrs, oracles, bs = systems(H)
    sizes = [A.rows for A in oracles]
    outs = []
    for graphs in (1, 0):
        S = H.Session(ctx, 4, sum(sizes[:4]), sum(len(a[1]) for a in arrs[:4]), params(H, "mixed", graphs))
        res, pending = {}, {}
        for k in range(4):  # staggered joins
            pending[S.add(*arrs[k], bs[k])] = k
            for slot in S.step():
                j = pending.pop(slot)
                res[j] = S.result(slot, sizes[j])
        run_to_end(S, pending, res, sizes)
        S.close()
        outs.append(res)
    for k in range(4):
        assert outs[0][k][1] == outs[1][k][1]
        np.testing.assert_array_equal(outs[0][k][0], outs[1][k][0])


def test_session_capacity_warm_start_and_errors(H, ctx):
    arrs, oracles, bs = systems(H)
    n0 = oracles[3].rows  # poisson 30^2
    p = params(H, "fp64")
    S = H.Session(ctx, 2, 2 * n0, 2 * len(arrs[3][1]), p)
    a = S.add(*arrs[3], bs[3])
    b = S.add(*arrs[3], 2 * bs[3])
    with pytest.raises(H.HspsError) as e:  # no slot left
        S.add(*arrs[3], bs[3])
    assert e.value.code == -6
    with pytest.raises(H.HspsError) as e:  # too large for the session at all
        S.add(*arrs[0], bs[0])
    assert e.value.code == -6
    res = {}
    run_to_end(S, {a: "a", b: "b"}, res, {"a": n0, "b": n0})
    xa, ita, cva, _ = res["a"]
    xb, itb, cvb, _ = res["b"]
    assert cva and cvb and ita == itb
    np.testing.assert_array_equal(2 * xa, xb)  # b scaled by 2: same iterates scaled by 2 (exact in binary)
    # warm start from the device copy of a finished solution: done at the initial check
    xd = ctx.array(xa)
    c = S.add(*arrs[3], bs[3], x0=xd)
    out = {}
    run_to_end(S, {c: "c"}, out, {"c": n0})
    x, it, cv, tr = out["c"]
    assert cv and it == 0 and tr < 1e-8
    np.testing.assert_array_equal(x, xa)
    # a zero diagonal is the Jacobi setup's error; the session stays usable
    rp, ci, v = arrs[3]
    v0 = v.copy()
    v0[np.flatnonzero(ci == 0)[0]] = 0.0
    with pytest.raises(H.HspsError) as e:
        S.add(rp, ci, v0, bs[3])
    assert e.value.code == -3
    with pytest.raises(H.HspsError):  # not a finished slot
        S.result(1, n0)
    assert S.live() == 0
    S.close()
