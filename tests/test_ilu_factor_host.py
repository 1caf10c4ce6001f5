"""The product's host ILU factorisation (hsps_ilu_factor: ILU(0) and the
sparse-scan ILUT) against the oracle's restatement of the reference's dense
row scan (ilu.cpp:41-89, 94-220), bit for bit, and against the reference
build's fixture.  Host code only: runs without a GPU."""
import numpy as np
import pytest

import oracle as O

CV, SV = 0.7648421872844885, 0.644217687237691


def _random(n, seed, dens=0.05):
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (n, n)) * (rng.uniform(size=(n, n)) < dens)
    D += np.diag(rng.uniform(1, 3, n) * np.sign(rng.uniform(-1, 1, n)))
    r, c = np.nonzero(D)
    return O.Csr.from_coo(n, n, r, c, D[r, c])


CASES = {
    "poisson30": lambda: O.Csr.poisson(30),
    "convdiff25": lambda: O.Csr.convdiff(25, 0.01, CV, SV),
    "convdiff20_strong": lambda: O.Csr.convdiff(20, 0.001, 1.0, 0.0),
    "aniso24": lambda: O.Csr.aniso(24, 0.01),
    "random300": lambda: _random(300, 4),
    "random120_dense": lambda: _random(120, 5, 0.3),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tau,fill", [(0.0, 0), (1e-4, 0), (1e-2, 0), (1e-3, 2), (0.1, 1)])
def test_ilu_factor_equals_oracle(H, case, tau, fill):
    A = CASES[case]()
    rp, ci, v = A.arrays()
    lrp, lci, lv = H.ilu_factor(rp, ci, v, tau, fill)
    orp, oci, ov = O.ilu_factor(A, drop_tol=tau, fill=fill).arrays()
    np.testing.assert_array_equal(lrp, orp)
    np.testing.assert_array_equal(lci, oci)
    np.testing.assert_array_equal(lv.view(np.int64), ov.view(np.int64))


def test_ilu_factor_reference_fixture(H, golden):
    f = {c["case"]: c for c in golden["ilu"]}
    for name, A, tau in [("poisson8_ilu0", O.Csr.poisson(8), 0.0),
                         ("convdiff12_ilut1e-4", O.Csr.convdiff(12, 0.1, 5.0, 0.0), 1e-4)]:
        lrp, lci, lv = H.ilu_factor(*A.arrays(), tau, 0)
        np.testing.assert_array_equal(lrp, f[name]["rp"])
        np.testing.assert_array_equal(lci, f[name]["ci"])
        np.testing.assert_allclose(lv, f[name]["lu"], rtol=1e-12)


def test_ilut_scales_to_serving_sizes(H):
    """ILUT(1e-4) of a 1024^2 convection-diffusion matrix: the dense scan is
    O(n^2) (10^12 steps here); the sparse scan finishes in seconds."""
    import time
    rp, ci, v = H.convection_diffusion_2d(1024, 0.01, CV, SV)
    t0 = time.perf_counter()
    lrp, lci, lv = H.ilu_factor(rp, ci, v, 1e-4, 0)
    dt = time.perf_counter() - t0
    assert len(lrp) == 1024 * 1024 + 1 and dt < 20.0, dt
    assert np.all(np.diff(lrp) >= 1)


def test_ilu_factor_zero_pivot_is_an_error(H):
    A = O.Csr.from_coo(2, 2, [0, 1], [1, 0], [1.0, 1.0])
    with pytest.raises(H.HspsError) as e:
        H.ilu_factor(*A.arrays(), 0.0, 0)
    assert e.value.code == -4
