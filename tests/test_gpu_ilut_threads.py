"""FGMRES + ILUT(1e-4) on a 160,000-row convection-diffusion matrix: the
preconditioner factored by the host ILUT on one thread and on its default
threads gives the same solve, bit for bit (the factors are the same bits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ilut_threads_same_solve(H, ctx, monkeypatch, precision):
    rp, ci, v = H.convection_diffusion_2d(400, 0.01, 0.6, 0.8)
    n = len(rp) - 1
    A = H.Csr(ctx, n, n, rp, ci, v)
    xs = ctx.array(H.uniform_pm1(5, n))
    b = ctx.array(n=n)
    A.spmv(xs, b)
    out = []
    for threads in ("1", None):
        if threads:
            monkeypatch.setenv("HSPS_ILUT_THREADS", threads)
        else:
            monkeypatch.delenv("HSPS_ILUT_THREADS", raising=False)
        p = H.default_params(tol=1e-9, max_iter=2000, restart_size=30, drop_tol=1e-4, precision=precision)
        M = H.Precond(A, "ilu", p)
        x = ctx.array(n=n)
        ok, st = H.solve(A, b, x, "fgmres", M, p)
        assert ok
        out.append((st.iterations, x.download()))
        M.close()
        x.free()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    A.close()
