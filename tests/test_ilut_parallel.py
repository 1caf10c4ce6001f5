"""The host ILUT on several threads (rows in blocks dealt round-robin; a row
waits only for the finished rows it eliminates with) gives the factors of the
one-thread scan bit for bit.  Host code only: runs without a GPU."""
import numpy as np
import pytest

import oracle as O

CV, SV = 0.7648421872844885, 0.644217687237691


def _banded(n, bw, seed):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        js = np.unique(np.clip(i + rng.integers(-bw, bw + 1, 6), 0, n - 1))
        for j in js:
            if j != i:
                rows.append(i); cols.append(j); vals.append(rng.uniform(-1, 1))
        rows.append(i); cols.append(i); vals.append(rng.uniform(3, 5))
    return O.Csr.from_coo(n, n, np.array(rows), np.array(cols), np.array(vals)).arrays()


CASES = {
    "convdiff300_pe100": lambda H: H.convection_diffusion_2d(300, 0.01, CV, SV),
    "convdiff200_strong": lambda H: H.convection_diffusion_2d(200, 0.001, 1.0, 0.0),
    "poisson256": lambda H: H.poisson_2d(256),
    "banded_random": lambda H: _banded(40000, 700, 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tau,fill", [(1e-4, 0), (1e-2, 3)])
def test_ilut_threads_give_the_same_factors(H, monkeypatch, case, tau, fill):
    rp, ci, v = CASES[case](H)
    out = []
    for t in (1, 2, 5, 8):
        monkeypatch.setenv("HSPS_ILUT_THREADS", str(t))
        out.append(H.ilu_factor(rp, ci, v, tau, fill))
    for lrp, lci, lv in out[1:]:
        np.testing.assert_array_equal(lrp, out[0][0])
        np.testing.assert_array_equal(lci, out[0][1])
        np.testing.assert_array_equal(lv.view(np.int64), out[0][2].view(np.int64))


def test_ilut_threads_small_cases_equal_oracle(H, monkeypatch):
    """Blocks of 512 rows on 3 threads over small matrices, against the
    oracle's restatement of the reference's dense scan."""
    monkeypatch.setenv("HSPS_ILUT_THREADS", "3")
    for A in (O.Csr.convdiff(40, 0.01, CV, SV), O.Csr.poisson(36)):
        rp, ci, v = A.arrays()
        lrp, lci, lv = H.ilu_factor(rp, ci, v, 1e-4, 0)
        orp, oci, ov = O.ilu_factor(A, drop_tol=1e-4, fill=0).arrays()
        np.testing.assert_array_equal(lrp, orp)
        np.testing.assert_array_equal(lci, oci)
        np.testing.assert_array_equal(lv.view(np.int64), ov.view(np.int64))
