"""Wavefront ILU sweeps over a banded grid layout (k_tri_band) against the
level-scheduled sweeps (HSPS_ILU_BAND=0 at setup) and the FP64 oracle, for
both band kernels (several workgroups with a global-memory handoff, the
default; one workgroup with LDS rings: HSPS_ILU_BAND=1).

Every row is the same expression in the same order on both paths, so the
bar is bit equality: the FP64 apply through the C ABI, and a mixed FGMRES
solve (FP32 sweeps inside) whose iterations and x agree bit for bit.  The
cases cover ILU(0) of the C3 geometry (1024^2 conv-diff, one round of 16
waves), 2048^2 (32 bands: two rounds), a grid whose last band is partial
(150^2), ILUT(1e-4) factors with fill (4 and 8 entry slots, lane skews up
to 4, an empty L), and the oracle's ILU apply within 1e-12.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CV, SV = 0.7648421872844885, 0.644217687237691
CASES = {
    "convdiff1024_ilu0": (lambda: O.Csr.convdiff(1024, 0.01, CV, SV), 0.0),
    "poisson2048_ilu0": (lambda: O.Csr.poisson(2048), 0.0),
    "poisson150_ilu0": (lambda: O.Csr.poisson(150), 0.0),
    "convdiff300_ilut": (lambda: O.Csr.convdiff(300, 0.05, 0.3, 0.95), 1e-4),
    "convdiff512_ilut_strong": (lambda: O.Csr.convdiff(512, 0.001, 1.0, 0.0), 1e-4),
    "aniso300_ilut": (lambda: O.Csr.aniso(300, 1.0), 1e-4),
}


def dev_csr(H, ctx, A):
    rp, ci, v = A.arrays()
    return H.Csr(ctx, A.rows, A.rows, rp, ci, v)


def two_preconds(H, ctx, monkeypatch, A, tau, mode="2"):
    """The band sweep (mode 2: several workgroups, 1: one workgroup) and the level sweep."""
    Ad = dev_csr(H, ctx, A)
    p = H.default_params(drop_tol=tau)
    monkeypatch.setenv("HSPS_ILU_BAND", mode)
    Mb = H.Precond(Ad, "ilu", p)
    monkeypatch.setenv("HSPS_ILU_BAND", "0")
    Ml = H.Precond(Ad, "ilu", p)
    monkeypatch.delenv("HSPS_ILU_BAND")
    return Ad, Mb, Ml


@pytest.mark.parametrize("mode", ["2", "1"])
@pytest.mark.parametrize("case", list(CASES))
def test_band_apply_equals_level_apply(H, ctx, monkeypatch, case, mode):
    make, tau = CASES[case]
    A = make()
    Ad, Mb, Ml = two_preconds(H, ctx, monkeypatch, A, tau, mode)
    (fk, fw), (bk, bw) = Mb.ilu_sweeps()
    assert bk == 1 and bw > 0, (case, Mb.ilu_sweeps())  # the backward sweep takes the band path
    assert Ml.ilu_sweeps() == ((0, 0), (0, 0))
    print(case, "sweeps", Mb.ilu_sweeps(), flush=True)
    r = O.uniform(17, A.rows)
    rd, zb, zl = ctx.array(r), ctx.array(n=A.rows), ctx.array(n=A.rows)
    Mb.apply(rd, zb)
    Ml.apply(rd, zl)
    z = zb.download()
    np.testing.assert_array_equal(z, zl.download())
    Mb.apply(rd, zb)  # a second apply reuses the skewed buffers
    np.testing.assert_array_equal(zb.download(), z)
    if A.rows <= 300 * 300:
        zo = O.precond_apply(A, r, "ilu", drop_tol=tau)
        assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)


@pytest.mark.parametrize("case", ["convdiff1024_ilu0", "convdiff300_ilut"])
def test_band_mixed_fgmres_equals_level(H, ctx, monkeypatch, case):
    """FP32 sweeps inside mixed FGMRES: same iterations, same x bits."""
    make, tau = CASES[case]
    A = make()
    Ad, Mb, Ml = two_preconds(H, ctx, monkeypatch, A, tau)
    b = A.spmv(O.uniform(23, A.rows))
    out = []
    for M in (Mb, Ml):
        p = H.default_params(tol=1e-10, max_iter=2000, restart_size=30, drop_tol=tau, precision="mixed")
        bd, xd = ctx.array(b), ctx.array(n=A.rows)
        ok, st = H.solve(Ad, bd, xd, "fgmres", M, p)
        assert ok
        out.append((st.iterations, xd.download()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
