"""CPU checks of the oracle's s-step CG, pipelined FGMRES and Lanczos spectral
interval (sstep_cg_solver.cpp, pipelined_fgmres_solver.cpp,
sparse_matrix.cpp:451-479).  The reference's tests hold no golden values for
s-step CG, so it is pinned by what the mathematics fixes: exact special cases
(identity, c*I through the Cholesky-failure fallback), the projection property
of one batch checked against an independent monomial Krylov basis, and k = 1
reducing to Richardson iteration.  Pipelined FGMRES is pinned by the
reference's own tests (test_improvements.cpp:224-283) and by equality with the
FGMRES restatement, itself pinned to the reference's iteration counts."""
import numpy as np

import oracle as O


def _diag(d):
    n = len(d)
    return O.Csr.from_coo(n, n, list(range(n)), list(range(n)), list(d))


def _dense(A):
    rp, ci, v = A.arrays()
    n, m, _ = A.shape_nnz()
    D = np.zeros((n, m))
    for i in range(n):
        D[i, ci[rp[i]:rp[i + 1]]] += v[rp[i]:rp[i + 1]]
    return D


def _scaled_poisson(side, f=0.25):
    rp, ci, v = O.Csr.poisson(side).arrays()
    n = side * side
    return O.Csr.from_csr(n, n, rp, ci, v * f)


def test_spectral_interval_exact_cases():
    assert O.spectral_interval(_diag([3.5])) == (3.5, 3.5)
    lo, hi = O.spectral_interval(_diag([2.0] * 5))
    assert abs(lo - 2.0) < 1e-15 and abs(hi - 2.0) < 1e-15
    # q = (1,1)/sqrt2 is an eigenvector of [[3,1],[1,3]] (eigenvalue 4)
    lo, hi = O.spectral_interval(O.Csr.from_coo(2, 2, [0, 0, 1, 1], [0, 1, 0, 1], [3.0, 1.0, 1.0, 3.0]))
    assert abs(lo - 4.0) < 1e-14 and abs(hi - 4.0) < 1e-14


def test_spectral_interval_encloses_the_spectrum():
    # n <= steps distinct eigenvalues: the Lanczos tridiagonal reproduces them,
    # and its Gershgorin bounds enclose them
    d = np.array([0.5, 1.0, 2.0, 3.0, 7.5, 9.0])
    lo, hi = O.spectral_interval(_diag(d))
    assert lo <= d.min() + 1e-12 and hi >= d.max() - 1e-12
    # Poisson 16: extreme eigenvalues 4 -+ 4 cos(pi/17)
    lo, hi = O.spectral_interval(O.Csr.poisson(16))
    lmax = 4 + 4 * np.cos(np.pi / 17)
    assert hi >= lmax - 1e-3 and lo < 4 - 4 * np.cos(np.pi / 17) + 0.5


def test_sstep_identity_and_scaled_identity():
    n = 40
    b = O.uniform(3, n)
    x, st = O.solve(_diag([1.0] * n), b, "sstep_cg", "none", sstep_k=1)
    assert st.converged and st.iterations == 1 and np.abs(x - b).max() < 1e-15
    # c*I with k = 4: the normalised Newton basis repeats +-B_0, the Gram
    # matrix is singular, Cholesky fails, and the fallback step is exact
    c = 2.5
    x, st = O.solve(_diag([c] * n), b, "sstep_cg", "none", sstep_k=4)
    assert st.converged and st.iterations == 1 and np.abs(x - b / c).max() < 1e-15


def test_sstep_batch_is_the_krylov_projection_of_r():
    """One batch: x1 - x0 is the orthogonal projection of r0 onto K_k(A, r0),
    checked with a QR of the monomial basis (not the Newton basis)."""
    A = _scaled_poisson(8)
    D = _dense(A)
    b = O.uniform(9, A.rows)
    for k in (2, 3, 4):
        x, st = O.solve(A, b, "sstep_cg", "none", sstep_k=k, max_iter=k)
        assert st.iterations == k
        K = np.empty((A.rows, k))
        K[:, 0] = b
        for j in range(1, k):
            K[:, j] = D @ K[:, j - 1]
        Q, _ = np.linalg.qr(K)
        proj = Q @ (Q.T @ b)
        assert np.linalg.norm(x - proj) <= 1e-9 * np.linalg.norm(proj), k


def test_sstep_k1_is_richardson():
    A = _scaled_poisson(12)
    D = _dense(A)
    b = O.uniform(5, A.rows)
    x, st = O.solve(A, b, "sstep_cg", "none", sstep_k=1, tol=1e-8, max_iter=20000)
    xr = np.zeros(A.rows)
    r = b.copy()
    it = 0
    while np.linalg.norm(r) / np.linalg.norm(b) > 1e-8:
        xr += r
        r = b - D @ xr
        it += 1
    assert st.converged and abs(st.iterations - it) <= 1
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


def test_sstep_flop_model():
    """Flops per batch: k SpMVs (basis) + k dots (B^T r) + 1 SpMV (residual)."""
    A = _scaled_poisson(10)
    n, _, nz = A.shape_nnz()
    k = 4
    x, st = O.solve(A, O.uniform(2, n), "sstep_cg", "none", sstep_k=k, max_iter=12)
    batches = st.iterations // k
    assert st.iterations == 12 and st.flop_count == batches * (k * 2 * nz + k * 2 * n + 2 * nz)


def test_pipelined_fgmres_reference_tests():
    """test_improvements.cpp:226-283."""
    x, st = O.solve(_diag([2.0, 4.0, 8.0]), np.array([4.0, 8.0, 16.0]), "pipelined_fgmres", "jacobi", tol=1e-10)
    np.testing.assert_allclose(x, [2.0, 2.0, 2.0], atol=1e-8)
    P = O.Csr.poisson(16)
    b = np.ones(256)
    x, st = O.solve(P, b, "pipelined_fgmres", "ilu", drop_tol=0.0, tol=1e-10)
    assert st.converged and np.linalg.norm(b - P.spmv(x)) / np.linalg.norm(b) < 1e-7
    P12 = O.Csr.poisson(12)
    b = np.ones(144)
    xs, _ = O.solve(P12, b, "fgmres", "ilu", drop_tol=0.0, tol=1e-10)
    xp, _ = O.solve(P12, b, "pipelined_fgmres", "ilu", drop_tol=0.0, tol=1e-10)
    assert np.linalg.norm(xs - xp) / np.linalg.norm(xs) < 1e-5
    Cd = O.Csr.convdiff(10, 0.1, 5.0, 0.0)
    x, st = O.solve(Cd, np.ones(100), "pipelined_fgmres", "ilu", drop_tol=0.0, tol=1e-10)
    assert st.converged


def test_pipelined_fgmres_equals_fgmres_without_breakdown():
    for A, pre in [(O.Csr.poisson(20), "ilu"), (O.Csr.convdiff(24, 0.02, 0.8, 0.6), "ilu"),
                   (O.Csr.poisson(24), "amg")]:
        n, _, nz = A.shape_nnz()
        b = O.uniform(4, n)
        xs, ss = O.solve(A, b, "fgmres", pre, drop_tol=0.0, restart_size=20, tol=1e-9)
        xp, sp = O.solve(A, b, "pipelined_fgmres", pre, drop_tol=0.0, restart_size=20, tol=1e-9)
        assert sp.iterations == ss.iterations and sp.restarts == ss.restarts
        assert np.linalg.norm(xs - xp) <= 1e-12 * np.linalg.norm(xs)
        cycles = sp.restarts + 1
        assert sp.flop_count == ss.flop_count + cycles * 2 * nz
