"""hsps_cg_solve_noalloc (PCG with its workspace in the library's static
per-device scratch array, no allocation) against hsps_cg_solve on the plain
kernel sequence (cg_engine = 0): the same kernels in the same order, so the
same bits.  Two calls in a row, the second on a smaller system (the scratch
keeps the first call's contents), for FP64 and mixed precision and for each
preconditioner kind; a system whose workspace exceeds the array is refused
before any device work."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(H):
    return H.lib().hsps_cg_solve_noalloc  # argument types from the binding's table


def _system(H, ctx, A, seed):
    rp, ci, v = A.arrays()
    n = A.rows
    b = A.spmv(O.uniform(seed, n))
    return H.Csr(ctx, n, n, rp, ci, v), b


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pre", ["none", "jacobi", "ilu", "amg"])
def test_noalloc_cg_equals_cg(H, ctx, precision, pre):
    f = _noalloc(H)
    # the second system is smaller than the first (rows and workspace)
    for k, A in enumerate([O.Csr.poisson(300), O.Csr.aniso(90, 0.05)]):
        Ad, b = _system(H, ctx, A, 30 + k)
        p = H.default_params(tol=1e-8, max_iter=20000, precision=precision, cg_engine=0, drop_tol=0.0)
        M = H.Precond(Ad, pre, p) if pre != "none" else None
        bd = ctx.array(b)
        x_ref, x_na = ctx.array(n=A.rows), ctx.array(n=A.rows)
        ok1, s1 = H.solve(Ad, bd, x_ref, "cg", M, p)
        s2 = H.Stats()
        rc = H.check(f(Ad.h, M.h if M else None, bd.ptr, x_na.ptr, C.byref(p), C.byref(s2)))
        print(f"{pre} {precision} n={A.rows}: {s1.iterations} / {s2.iterations} iterations", flush=True)
        assert ok1 and rc == 1
        assert (s2.iterations, s2.restarts, s2.reliable_updates) == (s1.iterations, s1.restarts, s1.reliable_updates)
        assert (s2.flop_count, s2.mem_bytes) == (s1.flop_count, s1.mem_bytes)
        assert s2.true_residual == s1.true_residual and s2.true_residual < 1e-8
        np.testing.assert_array_equal(x_na.download(), x_ref.download())
        if M:
            M.close()
        Ad.close()


def test_noalloc_cg_refuses_oversized_systems(H, ctx):
    """Unpreconditioned FP64 PCG needs 32 B of workspace per row (r, p, q
    and the FP64 residual): Poisson 1500^2 (72 MB) exceeds the 64 MiB array
    and is refused with HSPS_ERR_NOMEM before any device work."""
    f = _noalloc(H)
    side = 1500
    rp, ci, v = H.poisson_2d(side)
    n = side * side
    Ad = H.Csr(ctx, n, n, rp, ci, v)
    bd, xd = ctx.array(np.ones(n)), ctx.array(n=n)
    p = H.default_params(tol=1e-8, max_iter=10)
    s = H.Stats()
    rc = f(Ad.h, None, bd.ptr, xd.ptr, C.byref(p), C.byref(s))
    assert rc == -6 and b"scratch" in H.lib().hsps_last_error()
    Ad.close()
