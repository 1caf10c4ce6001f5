"""Pin the oracle's matrix generators by what the mathematics fixes.

* anisotropic_2d (no reference counterpart): -eps d_xx - d_yy on an n x n
  grid, DOF (i, j) -> i*n + j with x along the fast index j, the convention of
  convection_diffusion_2d (sparse_matrix.cpp:317-350, x-neighbours at j +- 1).
  Built independently here as eps * kron(I, T) + kron(T, I) with
  T = tridiag(-1, 2, -1), and checked entry by entry plus the closed forms:
  diagonal 2 eps + 2, j +- 1 -> -eps, i +- 1 -> -1, nnz = 5 N - 4 n,
  interior row sums 0.
* poisson_2d (sparse_matrix.cpp:288-311): kron(I, T) + kron(T, I).
* convection_diffusion_2d: the closed-form entries of sparse_matrix.cpp:317-350
  written from the formula (h = 1/(n+1), eh = eps/h^2, bxh = bx/(2h)).

A transposed operand (eps on the i +- 1 neighbours), a dropped boundary check
or a wrong diagonal fails one of these.
"""
import numpy as np
import pytest

import oracle as O


def _dense(A):
    rp, ci, v = A.arrays()
    n = len(rp) - 1
    D = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            D[i, ci[k]] += v[k]
    return D, rp, ci, v


def _tri(n):
    return 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)


@pytest.mark.parametrize("n,eps", [(1, 0.5), (2, 1.0), (5, 0.01), (12, 0.3), (17, 1e-3)])
def test_anisotropic_kron_closed_form(n, eps):
    A = O.Csr.aniso(n, eps)
    D, rp, ci, v = _dense(A)
    N = n * n
    K = eps * np.kron(np.eye(n), _tri(n)) + np.kron(_tri(n), np.eye(n))
    np.testing.assert_array_equal(D, K)
    assert len(v) == 5 * N - 4 * n
    # sorted, duplicate-free columns per row (the CSR invariant from_coo keeps)
    for i in range(N):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
    for i in range(n):
        for j in range(n):
            r = i * n + j
            assert D[r, r] == 2.0 * eps + 2.0
            if j > 0:
                assert D[r, r - 1] == -eps
            if j < n - 1:
                assert D[r, r + 1] == -eps
            if i > 0:
                assert D[r, r - n] == -1.0
            if i < n - 1:
                assert D[r, r + n] == -1.0
            if 0 < i < n - 1 and 0 < j < n - 1:  # zero up to the rounding of 2 eps + 2
                assert abs(D[r].sum()) <= 4 * np.finfo(float).eps * (2.0 * eps + 2.0)


def test_anisotropic_eps_one_is_poisson():
    np.testing.assert_array_equal(_dense(O.Csr.aniso(9, 1.0))[0], _dense(O.Csr.poisson(9))[0])


@pytest.mark.parametrize("n", [1, 3, 8])
def test_poisson_kron(n):
    D, _, _, v = _dense(O.Csr.poisson(n))
    np.testing.assert_array_equal(D, np.kron(np.eye(n), _tri(n)) + np.kron(_tri(n), np.eye(n)))
    assert len(v) == 5 * n * n - 4 * n


@pytest.mark.parametrize("n,eps,bx,by", [(6, 0.1, 5.0, 0.0), (7, 0.01, -0.7, 0.4), (5, 1.0, 0.3, -2.0)])
def test_convdiff_closed_form(n, eps, bx, by):
    D, _, _, v = _dense(O.Csr.convdiff(n, eps, bx, by))
    h = 1.0 / (n + 1)
    eh, bxh, byh = eps / (h * h), bx / (2 * h), by / (2 * h)
    for i in range(n):
        for j in range(n):
            r = i * n + j
            d = 4 * eh
            if j > 0:
                assert D[r, r - 1] == pytest.approx(-eh + max(bxh, 0.0), rel=1e-15)
                d += max(-bxh, 0.0)
            if j < n - 1:
                assert D[r, r + 1] == pytest.approx(-eh - min(bxh, 0.0), rel=1e-15)
                d += max(bxh, 0.0)
            if i > 0:
                assert D[r, r - n] == pytest.approx(-eh + max(byh, 0.0), rel=1e-15)
                d += max(-byh, 0.0)
            if i < n - 1:
                assert D[r, r + n] == pytest.approx(-eh - min(byh, 0.0), rel=1e-15)
                d += max(byh, 0.0)
            assert D[r, r] == pytest.approx(d, rel=1e-14)
            if 0 < i < n - 1 and 0 < j < n - 1:  # the reference's inconsistent upwind row sum
                assert D[r].sum() == pytest.approx(2 * abs(bxh) + 2 * abs(byh), rel=1e-12, abs=1e-9 * eh)
    assert len(v) == 5 * n * n - 4 * n
