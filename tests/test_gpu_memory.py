"""The device memory cache (core.hip): freed blocks are reused once the
freeing context's stream has passed the free, and freeing never waits for
work queued on other contexts' streams (plain hipFree waits for every stream
on the device)."""
import ctypes as C
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _malloc(H, ctx, nbytes):
    p = C.c_void_p()
    H.check(H.lib().hsps_malloc(ctx.h, nbytes, C.byref(p)))
    return p


def test_freed_block_is_reused_by_its_size_class(H):
    """hsps_free drains the context's stream and parks the block in the idle
    pool; a context that owns no cached block of that size class takes the
    last block parked there (a fresh context: the earlier tests' contexts own
    blocks of their own)."""
    c = H.Context(0)
    try:
        p1 = _malloc(H, c, 3 << 23)
        H.check(H.lib().hsps_free(c.h, p1))
        p2 = _malloc(H, c, (3 << 23) - 4096)  # same class (3 * 2^23): the drained block comes back
        assert p2.value == p1.value
        p3 = _malloc(H, c, 1 << 25)  # another class
        assert p3.value != p2.value
        H.check(H.lib().hsps_free(c.h, p2))
        H.check(H.lib().hsps_free(c.h, p3))
    finally:
        c.close()


def test_free_does_not_wait_for_another_contexts_stream(H, ctx):
    other = H.Context(0)
    try:
        rp, ci, v = H.poisson_2d(2048)
        n = len(rp) - 1
        A = H.Csr(other, n, n, rp, ci, v)
        x, y = other.array(np.ones(n)), other.array(n=n)
        other.sync()
        for _ in range(400):  # ~58 us each: about 23 ms queued on the other context's stream
            A.spmv(x, y)
        t0 = time.perf_counter()
        p = _malloc(H, ctx, 64 << 20)
        H.check(H.lib().hsps_free(ctx.h, p))
        t_free = time.perf_counter() - t0
        t1 = time.perf_counter()
        other.sync()
        t_rest = time.perf_counter() - t1
        # the other stream was still busy after the free returned, and the free did not wait for it
        assert t_rest > 5e-3, (t_free, t_rest)
        assert t_free < 5e-3, (t_free, t_rest)
        np.testing.assert_allclose(y.download()[:4], [2.0, 1.0, 1.0, 1.0])
        x.free()
        y.free()
        A.close()
    finally:
        other.close()


def test_solve_on_reused_blocks_gives_the_same_bits(H, ctx):
    """A system solved, released (its buffers back in the cache, stale
    contents and all) and solved again on a fresh context gives identical
    iterates: nothing depends on freshly allocated memory being zero."""
    A0 = O.Csr.poisson(300)
    rp, ci, v = A0.arrays()
    n = A0.rows
    b = A0.spmv(O.uniform(17, n))
    xs = []
    for c in (ctx, H.Context(0)):
        for prec in ("fp64", "mixed"):
            A = H.Csr(c, n, n, rp, ci, v)
            p = H.default_params(tol=1e-8, max_iter=5000, precision=prec, cg_engine=0)
            M = H.Precond(A, "jacobi", p)
            bd, xd = c.array(b), c.array(n=n)
            ok, st = H.solve(A, bd, xd, "cg", M, p)
            assert ok
            xs.append((prec, st.iterations, xd.download()))
            M.close()
            A.close()
            bd.free()
            xd.free()
        if c is not ctx:
            c.close()
    for k in range(2):
        assert xs[k][1] == xs[k + 2][1]
        np.testing.assert_array_equal(xs[k][2], xs[k + 2][2])
