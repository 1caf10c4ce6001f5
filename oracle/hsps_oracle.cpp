// =============================================================================
// hsps_oracle.cpp — FP64 CPU oracle (test infrastructure only; see header).
// Citations are file:line in the reference repository.
// =============================================================================
#include "hsps_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

#include <omp.h>

namespace orc {

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---- from_coo: sparse_matrix.cpp:242-282 --------------------------------------
// Sort by (row, col); duplicate (row, col) pairs accumulate.
Csr from_coo(Index rows, Index cols, const std::vector<Index>& r,
             const std::vector<Index>& c, const std::vector<Real>& v) {
    const size_t nz = r.size();
    std::vector<size_t> ord(nz);
    std::iota(ord.begin(), ord.end(), size_t(0));
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
    });
    Csr A;
    A.rows = rows;
    A.cols = cols;
    A.rp.assign(rows + 1, 0);
    A.ci.reserve(nz);
    A.v.reserve(nz);
    Index last_r = -1, last_c = -1;
    for (size_t p = 0; p < nz; ++p) {
        const size_t o = ord[p];
        if (r[o] == last_r && c[o] == last_c) {
            A.v.back() += v[o];
        } else {
            A.ci.push_back(c[o]);
            A.v.push_back(v[o]);
            A.rp[r[o] + 1]++;
            last_r = r[o];
            last_c = c[o];
        }
    }
    for (Index i = 0; i < rows; ++i) A.rp[i + 1] += A.rp[i];
    return A;
}

// ---- poisson_2d: sparse_matrix.cpp:288-311 -------------------------------------
// 5-point, diagonal 4, neighbours -1, unscaled; DOF (i,j) -> i*n+j.
Csr poisson_2d(Index n) {
    const Index N = n * n;
    std::vector<Index> r, c;
    std::vector<Real> v;
    r.reserve(5 * (size_t)N); c.reserve(5 * (size_t)N); v.reserve(5 * (size_t)N);
    auto push = [&](Index a, Index b, Real x) { r.push_back(a); c.push_back(b); v.push_back(x); };
    for (Index i = 0; i < n; ++i)
        for (Index j = 0; j < n; ++j) {
            const Index row = i * n + j;
            push(row, row, 4.0);
            if (i > 0) push(row, (i - 1) * n + j, -1.0);
            if (i < n - 1) push(row, (i + 1) * n + j, -1.0);
            if (j > 0) push(row, i * n + j - 1, -1.0);
            if (j < n - 1) push(row, i * n + j + 1, -1.0);
        }
    return from_coo(N, N, r, c, v);
}

// ---- convection_diffusion_2d: sparse_matrix.cpp:317-350 -------------------------
Csr convection_diffusion_2d(Index n, Real eps, Real bx, Real by) {
    const Index N = n * n;
    const Real h = 1.0 / (n + 1);
    const Real eh = eps / (h * h);
    const Real bxh = bx / (2.0 * h);
    const Real byh = by / (2.0 * h);
    std::vector<Index> r, c;
    std::vector<Real> v;
    r.reserve(5 * (size_t)N); c.reserve(5 * (size_t)N); v.reserve(5 * (size_t)N);
    auto push = [&](Index a, Index b, Real x) { r.push_back(a); c.push_back(b); v.push_back(x); };
    for (Index i = 0; i < n; ++i)
        for (Index j = 0; j < n; ++j) {
            const Index row = i * n + j;
            Real diag = 4.0 * eh;
            if (j > 0) { push(row, i * n + j - 1, -eh + std::max(bxh, 0.0)); diag += std::max(-bxh, 0.0); }
            if (j < n - 1) { push(row, i * n + j + 1, -eh - std::min(bxh, 0.0)); diag += std::max(bxh, 0.0); }
            if (i > 0) { push(row, (i - 1) * n + j, -eh + std::max(byh, 0.0)); diag += std::max(-byh, 0.0); }
            if (i < n - 1) { push(row, (i + 1) * n + j, -eh - std::min(byh, 0.0)); diag += std::max(byh, 0.0); }
            push(row, row, diag);
        }
    return from_coo(N, N, r, c, v);
}

// Anisotropic diffusion (no reference counterpart).
Csr anisotropic_2d(Index n, Real eps) {
    const Index N = n * n;
    std::vector<Index> r, c;
    std::vector<Real> v;
    auto push = [&](Index a, Index b, Real x) { r.push_back(a); c.push_back(b); v.push_back(x); };
    for (Index i = 0; i < n; ++i)
        for (Index j = 0; j < n; ++j) {
            const Index row = i * n + j;
            push(row, row, 2.0 * eps + 2.0);
            if (i > 0) push(row, (i - 1) * n + j, -1.0);
            if (i < n - 1) push(row, (i + 1) * n + j, -1.0);
            if (j > 0) push(row, i * n + j - 1, -eps);
            if (j < n - 1) push(row, i * n + j + 1, -eps);
        }
    return from_coo(N, N, r, c, v);
}

// ---- spmv: sparse_matrix.cpp:68-87 (OpenMP dynamic,64 rows) ---------------------
void spmv(const Csr& A, const Real* x, Real* y) {
    const Index* rp = A.rp.data();
    const Index* ci = A.ci.data();
    const Real* val = A.v.data();
#pragma omp parallel for schedule(dynamic, 64)
    for (Index i = 0; i < A.rows; ++i) {
        Real s = 0.0;
        for (Index k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[ci[k]];
        y[i] = s;
    }
}

// ---- dot / norm2: vector.cpp:39-53 (OpenMP static reduction) -------------------
// The same static partition of the rows over the threads as the reference's
// reduction(+); the per-thread partial sums are then added in thread order
// (libgomp combines a reduction's partials in arrival order, which made the
// last bits, and the iteration counts of ill-conditioned solves, vary from
// run to run).
Real dot(Index n, const Real* a, const Real* b) {
    const int nt = omp_get_max_threads();
    std::vector<Real> part((size_t)nt, 0.0);
    int used = 1;
#pragma omp parallel
    {
        Real s = 0.0;
#pragma omp for schedule(static)
        for (Index i = 0; i < n; ++i) s += a[i] * b[i];
        part[(size_t)omp_get_thread_num()] = s;
#pragma omp single
        used = omp_get_num_threads();
    }
    Real s = 0.0;
    for (int t = 0; t < used; ++t) s += part[(size_t)t];
    return s;
}
Real norm2(Index n, const Real* a) { return std::sqrt(dot(n, a, a)); }

static void axpy(Index n, Real alpha, const Real* x, Real* y) {  // vector.cpp:73-80
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) y[i] += alpha * x[i];
}
static void axpby(Index n, Real a, const Real* x, Real b, Real* y) {  // vector.cpp:82-89
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) y[i] = a * x[i] + b * y[i];
}
static void scale(Index n, Real a, Real* y) {  // vector.cpp:66-71
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) y[i] *= a;
}

// ---- transpose: sparse_matrix.cpp:138-165 ---------------------------------------
Csr transpose(const Csr& A) {
    Csr T;
    T.rows = A.cols;
    T.cols = A.rows;
    T.rp.assign(A.cols + 1, 0);
    for (Index k = 0; k < A.nnz(); ++k) T.rp[A.ci[k] + 1]++;
    for (Index j = 0; j < A.cols; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(A.nnz());
    T.v.resize(A.nnz());
    std::vector<Index> pos(T.rp.begin(), T.rp.end() - 1);
    for (Index i = 0; i < A.rows; ++i)
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const Index p = pos[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

// ---- mat_mat_product: sparse_matrix.cpp:171-237 ---------------------------------
// Row-wise dense accumulator in (A row order, B row order); the reference
// build contracts `dense[l] += av*bv` into an FMA, so this does too.  Output
// columns sorted; entries with |v| <= eps^2 dropped.
Csr mat_mat_product(const Csr& A, const Csr& B) {
    if (B.rows != A.cols) throw std::runtime_error("mat_mat_product: inner dimensions mismatch");
    const Index m = A.rows, n = B.cols;
    std::vector<std::vector<std::pair<Index, Real>>> rows(m);
#pragma omp parallel
    {
        std::vector<Real> dense(n, 0.0);
        std::vector<Index> touched;
#pragma omp for schedule(dynamic, 64)
        for (Index i = 0; i < m; ++i) {
            for (Index pa = A.rp[i]; pa < A.rp[i + 1]; ++pa) {
                const Index j = A.ci[pa];
                const Real av = A.v[pa];
                for (Index pb = B.rp[j]; pb < B.rp[j + 1]; ++pb) {
                    const Index l = B.ci[pb];
                    if (dense[l] == 0.0) touched.push_back(l);
                    dense[l] = std::fma(av, B.v[pb], dense[l]);
                }
            }
            std::sort(touched.begin(), touched.end());
            for (Index l : touched) {
                if (std::abs(dense[l]) > EPS * EPS) rows[i].emplace_back(l, dense[l]);
                dense[l] = 0.0;
            }
            touched.clear();
        }
    }
    Csr C;
    C.rows = m;
    C.cols = n;
    C.rp.assign(m + 1, 0);
    for (Index i = 0; i < m; ++i) C.rp[i + 1] = C.rp[i] + (Index)rows[i].size();
    C.ci.resize(C.rp[m]);
    C.v.resize(C.rp[m]);
    for (Index i = 0; i < m; ++i) {
        Index p = C.rp[i];
        for (auto& e : rows[i]) { C.ci[p] = e.first; C.v[p] = e.second; ++p; }
    }
    return C;
}

// ---- extract_diagonal: sparse_matrix.cpp:98-109 ---------------------------------
std::vector<Real> diagonal(const Csr& A) {
    std::vector<Real> d(A.rows, 0.0);
    for (Index i = 0; i < A.rows; ++i)
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) { d[i] = A.v[k]; break; }
    return d;
}

static Real get(const Csr& A, Index i, Index j) {  // sparse_matrix.cpp:38-55
    auto b = A.ci.begin() + A.rp[i], e = A.ci.begin() + A.rp[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return 0.0;
    return A.v[it - A.ci.begin()];
}

// ---- seeded inputs ---------------------------------------------------------
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// x_i = 2*u - 1 with u = top 53 bits / 2^53 of the i-th splitmix64 draw.
void uniform_pm1(uint64_t seed, Index n, Real* out) {
    uint64_t s = seed;
    for (Index i = 0; i < n; ++i) {
        const double u = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
        out[i] = 2.0 * u - 1.0;
    }
}

// ---- Jacobi: jacobi.cpp:13-51 ------------------------------------------------
void Jacobi::setup(const Csr& A) {
    n = A.rows;
    inv_diag.assign(n, 0.0);
    for (Index i = 0; i < n; ++i) {
        Real d = 0.0;
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) { d = A.v[k]; break; }
        if (std::abs(d) < EPS)
            throw std::runtime_error("JacobiPreconditioner: zero diagonal at row " + std::to_string(i));
        inv_diag[i] = omega / d;
    }
}
void Jacobi::apply(const Real* r, Real* z) const {
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
}

// ---- ILU: ilu.cpp:31-271 -------------------------------------------------------
void Ilu::setup(const Csr& A) {  // ilu.cpp:31-36
    if (drop_tol > 0.0 || fill_p > 0) setup_ilut(A); else setup_ilu0(A);
}

void Ilu::setup_ilu0(const Csr& A) {  // ilu.cpp:41-89 (IKJ on A's pattern)
    n = A.rows;
    rp = A.rp;
    ci = A.ci;
    lu = A.v;
    std::vector<Index> pos(n, -1);
    for (Index i = 1; i < n; ++i) {
        for (Index p = rp[i]; p < rp[i + 1]; ++p) pos[ci[p]] = p;
        for (Index p = rp[i]; p < rp[i + 1]; ++p) {
            const Index k = ci[p];
            if (k >= i) break;
            Real dk = 0.0;
            for (Index q = rp[k]; q < rp[k + 1]; ++q)
                if (ci[q] == k) { dk = lu[q]; break; }
            if (std::abs(dk) < EPS)
                throw std::runtime_error("ILU(0): near-zero pivot at row " + std::to_string(k));
            lu[p] /= dk;
            const Real lik = lu[p];
            for (Index q = rp[k]; q < rp[k + 1]; ++q) {
                const Index j = ci[q];
                if (j <= k) continue;
                const Index pij = pos[j];
                if (pij < 0) continue;
                lu[pij] -= lik * lu[q];
            }
        }
        for (Index p = rp[i]; p < rp[i + 1]; ++p) pos[ci[p]] = -1;
    }
}

void Ilu::setup_ilut(const Csr& A) {  // ilu.cpp:94-220
    n = A.rows;
    std::vector<std::vector<std::pair<Index, Real>>> L(n), U(n);
    std::vector<Real> w(n, 0.0);
    std::vector<char> act(n, 0);
    for (Index i = 0; i < n; ++i) {
        Real rn = 0.0;
        for (Index p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const Index j = A.ci[p];
            w[j] = A.v[p];
            act[j] = 1;
            rn += A.v[p] * A.v[p];
        }
        rn = std::sqrt(rn);
        const Real tau = drop_tol * rn;
        for (Index k = 0; k < i; ++k) {
            if (!act[k]) continue;
            if (std::abs(w[k]) < tau) { w[k] = 0.0; act[k] = 0; continue; }
            Real ukk = 0.0;
            for (auto& e : U[k]) if (e.first == k) { ukk = e.second; break; }
            if (std::abs(ukk) < EPS) { act[k] = 0; continue; }
            const Real lik = w[k] / ukk;
            w[k] = lik;
            for (auto& e : U[k]) {
                const Index j = e.first;
                if (j <= k) continue;
                if (!act[j] && std::abs(lik * e.second) >= tau) act[j] = 1;
                w[j] -= lik * e.second;
            }
        }
        std::vector<std::pair<Index, Real>> Lr, Ur;
        for (Index j = 0; j < i; ++j) {
            if (act[j] && std::abs(w[j]) >= tau) Lr.emplace_back(j, w[j]);
            w[j] = 0.0; act[j] = 0;
        }
        Real d = w[i];
        if (std::abs(d) < EPS) d = (d >= 0 ? 1.0 : -1.0) * EPS * 1e6;
        Ur.emplace_back(i, d);
        w[i] = 0.0; act[i] = 0;
        for (Index j = i + 1; j < n; ++j) {
            if (act[j] && std::abs(w[j]) >= tau) Ur.emplace_back(j, w[j]);
            w[j] = 0.0; act[j] = 0;
        }
        auto keep = [](std::vector<std::pair<Index, Real>>& row, int mk) {
            if (mk <= 0 || (int)row.size() <= mk) return;
            std::partial_sort(row.begin(), row.begin() + mk, row.end(),
                              [](const auto& a, const auto& b) { return std::abs(a.second) > std::abs(b.second); });
            row.resize(mk);
            std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        };
        const int orig = A.rp[i + 1] - A.rp[i];
        keep(Lr, orig + fill_p);
        keep(Ur, orig + fill_p);
        L[i] = std::move(Lr);
        U[i] = std::move(Ur);
    }
    rp.assign(n + 1, 0);
    ci.clear();
    lu.clear();
    for (Index i = 0; i < n; ++i) {
        for (auto& e : L[i]) { ci.push_back(e.first); lu.push_back(e.second); }
        for (auto& e : U[i]) { ci.push_back(e.first); lu.push_back(e.second); }
        rp[i + 1] = (Index)ci.size();
    }
}

void Ilu::forward_sub(const Real* r, Real* y) const {  // ilu.cpp:225-239
    for (Index i = 0; i < n; ++i) {
        Real s = r[i];
        for (Index p = rp[i]; p < rp[i + 1]; ++p) {
            const Index j = ci[p];
            if (j >= i) break;
            s -= lu[p] * y[j];
        }
        y[i] = s;
    }
}

void Ilu::backward_sub(const Real* y, Real* z) const {  // ilu.cpp:244-262
    for (Index i = n - 1; i >= 0; --i) {
        Real s = y[i], d = 1.0;
        for (Index p = rp[i]; p < rp[i + 1]; ++p) {
            const Index j = ci[p];
            if (j == i) { d = lu[p]; continue; }
            if (j > i) s -= lu[p] * z[j];
        }
        if (std::abs(d) < EPS)
            throw std::runtime_error("ILU backward sub: near-zero U diagonal at row " + std::to_string(i));
        z[i] = s / d;
    }
}

void Ilu::apply(const Real* r, Real* z) const {  // ilu.cpp:267-271
    std::vector<Real> y(n);
    forward_sub(r, y.data());
    backward_sub(y.data(), z);
}

// ---- AMG: amg.cpp:30-343 -------------------------------------------------------
// Greedy aggregation, amg.cpp:30-67.  The singleton pass at :63-64 can never
// fire (every visited node is assigned), so it is not restated.
std::vector<Index> aggregate(const Csr& A, Real theta) {
    const Index n = A.rows;
    std::vector<Index> agg(n, -1);
    std::vector<Real> row_max(n, 0.0);
    for (Index i = 0; i < n; ++i)
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] != i) row_max[i] = std::max(row_max[i], std::abs(A.v[k]));
    Index nagg = 0;
    for (Index i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = nagg;
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const Index j = A.ci[k];
            if (j == i || agg[j] >= 0) continue;
            if (std::abs(A.v[k]) >= theta * row_max[i]) agg[j] = nagg;
        }
        ++nagg;
    }
    return agg;
}

Csr build_prolongation(Index nf, const std::vector<Index>& agg, Index nc) {  // amg.cpp:73-89
    Csr P;
    P.rows = nf;
    P.cols = nc;
    P.rp.resize(nf + 1);
    P.ci.resize(nf);
    P.v.assign(nf, 1.0);
    for (Index i = 0; i < nf; ++i) { P.rp[i] = i; P.ci[i] = agg[i]; }
    P.rp[nf] = nf;
    return P;
}

// smooth_prolongation_op, amg.cpp:101-178, restated sparsely.  The reference
// works in a dense nf x nc buffer; for one smoothing step on a one-entry-per-
// row P_raw, entry (i,c) is P_raw(i,c) minus, in row order over j != i,
// (omega/a_ii)*a_ij*P_raw(j,c).  Only columns reached from row i can be
// nonzero, so a per-row sparse accumulator gives the same values in the same
// order.  The diagonal of A is skipped; |v| <= 1e-14 is dropped; columns are
// emitted in ascending order.
Csr smooth_prolongation(const Csr& A, const Csr& Praw, Real omega) {
    const Index nf = A.rows, nc = Praw.cols;
    std::vector<Real> inv_diag(nf, 0.0);
    for (Index i = 0; i < nf; ++i)
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) { inv_diag[i] = std::abs(A.v[k]) > EPS ? omega / A.v[k] : 0.0; break; }
    Csr P;
    P.rows = nf;
    P.cols = nc;
    P.rp.assign(nf + 1, 0);
    std::vector<std::pair<Index, Real>> row;
    for (Index i = 0; i < nf; ++i) {
        row.clear();
        for (Index q = Praw.rp[i]; q < Praw.rp[i + 1]; ++q) row.emplace_back(Praw.ci[q], Praw.v[q]);
        const Real id = inv_diag[i];
        if (!(std::abs(id) < EPS)) {
            for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const Index j = A.ci[k];
                if (j == i) continue;
                const Real t = id * A.v[k];
                for (Index q = Praw.rp[j]; q < Praw.rp[j + 1]; ++q) {
                    const Index c = Praw.ci[q];
                    auto it = std::find_if(row.begin(), row.end(), [c](const auto& e) { return e.first == c; });
                    if (it == row.end()) { row.emplace_back(c, 0.0); it = row.end() - 1; }
                    it->second -= t * Praw.v[q];
                }
            }
        }
        std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& e : row)
            if (std::abs(e.second) > 1e-14) { P.ci.push_back(e.first); P.v.push_back(e.second); }
        P.rp[i + 1] = (Index)P.ci.size();
    }
    return P;
}

void Amg::setup(const Csr& A) {  // amg.cpp:183-237
    h.clear();
    n = A.rows;
    auto inv = [](const Csr& M) {
        std::vector<Real> d = diagonal(M);
        for (auto& x : d) x = std::abs(x) > EPS ? 1.0 / x : 0.0;
        return d;
    };
    AmgLevel l0;
    l0.A = A;
    l0.dinv = inv(A);
    h.push_back(std::move(l0));
    for (int lv = 0; lv < max_levels - 1; ++lv) {
        const Csr& Af = h.back().A;
        if (Af.rows <= coarse_limit) break;
        std::vector<Index> agg = aggregate(Af, theta);
        const Index nc = *std::max_element(agg.begin(), agg.end()) + 1;
        if (nc >= Af.rows) break;
        Csr Praw = build_prolongation(Af.rows, agg, nc);
        Csr P = smooth_p ? smooth_prolongation(Af, Praw, omega) : Praw;
        Csr R = transpose(P);
        Csr Ac = mat_mat_product(R, mat_mat_product(Af, P));
        AmgLevel l;
        l.A = std::move(Ac);
        l.dinv = inv(l.A);
        h.back().P = std::move(P);
        h.back().R = std::move(R);
        h.back().agg = std::move(agg);
        h.push_back(std::move(l));
    }
}

void Amg::smooth(int lv, const Real* rhs, Real* x, int steps) const {  // amg.cpp:242-254
    const Csr& A = h[lv].A;
    const auto& dinv = h[lv].dinv;
    std::vector<Real> res(A.rows);
    for (int s = 0; s < steps; ++s) {
        spmv(A, x, res.data());
        for (Index i = 0; i < A.rows; ++i) x[i] += omega * dinv[i] * (rhs[i] - res[i]);
    }
}

// direct_solve, amg.cpp:260-291: dense GEPP on [A | rhs], rebuilt per call.
void dense_gepp_solve(const Csr& A, const Real* rhs, Real* x) {
    const Index n = A.rows;
    std::vector<std::vector<Real>> M(n, std::vector<Real>(n + 1, 0.0));
    for (Index i = 0; i < n; ++i) {
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) M[i][A.ci[k]] = A.v[k];
        M[i][n] = rhs[i];
    }
    for (Index col = 0; col < n; ++col) {
        Index piv = col;
        for (Index r = col + 1; r < n; ++r)
            if (std::abs(M[r][col]) > std::abs(M[piv][col])) piv = r;
        std::swap(M[col], M[piv]);
        if (std::abs(M[col][col]) < 1e-14) continue;
        const Real ip = 1.0 / M[col][col];
        for (Index r = col + 1; r < n; ++r) {
            const Real f = M[r][col] * ip;
            for (Index c = col; c <= n; ++c) M[r][c] -= f * M[col][c];
        }
    }
    for (Index i = n - 1; i >= 0; --i) {
        Real s = M[i][n];
        for (Index j = i + 1; j < n; ++j) s -= M[i][j] * x[j];
        x[i] = std::abs(M[i][i]) > 1e-14 ? s / M[i][i] : 0.0;
    }
}

void Amg::v_cycle(int lv, const Real* rhs, Real* x) const {  // amg.cpp:296-335
    const int L = (int)h.size();
    if (lv == L - 1) { dense_gepp_solve(h[lv].A, rhs, x); return; }
    const Csr& A = h[lv].A;
    smooth(lv, rhs, x, pre);
    std::vector<Real> ax(A.rows), res(A.rows);
    spmv(A, x, ax.data());
    for (Index i = 0; i < A.rows; ++i) res[i] = rhs[i] - ax[i];
    std::vector<Real> rc(h[lv].R.rows), ec(h[lv + 1].A.rows, 0.0), ef(A.rows);
    spmv(h[lv].R, res.data(), rc.data());
    v_cycle(lv + 1, rc.data(), ec.data());
    spmv(h[lv].P, ec.data(), ef.data());
    for (Index i = 0; i < A.rows; ++i) x[i] += ef[i];
    smooth(lv, rhs, x, post);
}

void Amg::apply(const Real* r, Real* z) const {  // amg.cpp:340-343 (z zeroed)
    std::fill(z, z + n, 0.0);
    v_cycle(0, r, z);
}

// ---- flop_model: energy_monitor.hpp:163-177 --------------------------------------
static long long spmv_flops(long long nnz) { return 2LL * nnz; }
static long long spmv_bytes(long long rows, long long nnz) {
    return (rows + 1 + nnz) * 4 + nnz * 8 + rows * 8 * 2;
}

// ---- CG: cg_solver.cpp:37-191 ----------------------------------------------------
bool cg_solve(const Csr& A, const Real* b, Real* x, const Precond* M, const Params& prm, Stats& st) {
    const double t0 = now_s();
    const Index n = A.rows;
    const long long nz = A.nnz();
    std::vector<Real> r(n), z(n), p(n), q(n);
    spmv(A, x, r.data());
    for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const Real bn = norm2(n, b), r0 = norm2(n, r.data());
    st.initial_residual = r0;
    if (bn < EPS) { st.converged = true; st.final_residual = r0; st.iterations = 0; return true; }
    Real rel = r0 / bn;
    if (rel < prm.tol) { st.converged = true; st.final_residual = rel; st.iterations = 0; return true; }
    auto prec = [&](const Real* in, Real* out) {
        if (M) M->apply(in, out); else std::copy(in, in + n, out);
    };
    prec(r.data(), z.data());
    p = z;
    Real rho = dot(n, r.data(), z.data()), rho_old = rho;
    long long F = 0, B = 0;
    int it = 0;
    for (; it < prm.max_iter; ++it) {
        spmv(A, p.data(), q.data());
        F += spmv_flops(nz); B += spmv_bytes(n, nz);
        const Real pq = dot(n, p.data(), q.data());
        F += 2LL * n; B += 16LL * n;
        if (std::abs(pq) < EPS * 1e6 * std::abs(rho)) {  // :109-120
            spmv(A, x, p.data());
            for (Index i = 0; i < n; ++i) r[i] = b[i] - p[i];
            prec(r.data(), z.data());
            p = z;
            rho = dot(n, r.data(), z.data());
            ++st.restarts;
            continue;
        }
        const Real alpha = rho / pq;
        axpy(n, alpha, p.data(), x);
        axpy(n, -alpha, q.data(), r.data());
        F += 4LL * n; B += 48LL * n;
        rel = norm2(n, r.data()) / bn;
        if (rel < prm.tol) { ++it; break; }
        prec(r.data(), z.data());
        rho_old = rho;
        rho = dot(n, r.data(), z.data());
        F += 2LL * n; B += 16LL * n;
        if (std::abs(rho_old) < EPS * 1e8 * std::abs(rho)) {  // :157-167
            spmv(A, x, z.data());
            for (Index i = 0; i < n; ++i) r[i] = b[i] - z[i];
            prec(r.data(), z.data());
            p = z;
            rho = dot(n, r.data(), z.data());
            rho_old = rho;
            ++st.restarts;
            continue;
        }
        const Real beta = rho / rho_old;
        axpby(n, 1.0, z.data(), beta, p.data());
        F += 2LL * n; B += 24LL * n;
    }
    st.iterations = it;
    st.final_residual = rel;
    st.converged = rel < prm.tol;
    st.solve_time_s = now_s() - t0;
    st.flop_count = F;
    st.mem_bytes = B;
    st.energy_joules = 2.0e-10 * F + 5.0e-9 * B;
    return st.converged;
}

// ---- FGMRES: fgmres_solver.cpp:36-241 -----------------------------------------------
static void givens(Real f, Real g, Real& cs, Real& sn, Real& r) {  // :36-45
    if (std::abs(g) < EPS) { cs = 1.0; sn = 0.0; r = f; }
    else if (std::abs(f) < EPS) { cs = 0.0; sn = 1.0; r = g; }
    else { const Real t = std::sqrt(f * f + g * g); cs = f / t; sn = g / t; r = t; }
}
static void rot(Real& dx, Real& dy, Real cs, Real sn) {  // :47-51
    const Real t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

static int fgmres_cycle(const Csr& A, const Real* b, Real* x, const Precond* M, Real tol_abs, int m,
                        Real& res, std::vector<std::vector<Real>>& V, std::vector<std::vector<Real>>& Z) {
    const Index n = A.rows;
    std::vector<Real> r(n), w(n);
    spmv(A, x, r.data());
    for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const Real beta = norm2(n, r.data());
    res = beta;
    if (beta < tol_abs) return 0;
    V.assign(m + 1, std::vector<Real>(n, 0.0));
    Z.assign(m, std::vector<Real>(n, 0.0));
    V[0] = r;
    scale(n, 1.0 / beta, V[0].data());
    std::vector<std::vector<Real>> H(m + 1, std::vector<Real>(m, 0.0));
    std::vector<Real> cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m; ++j) {
        if (M) M->apply(V[j].data(), Z[j].data()); else Z[j] = V[j];
        spmv(A, Z[j].data(), w.data());
        for (int i = 0; i <= j; ++i) {  // MGS, :122-125
            H[i][j] = dot(n, w.data(), V[i].data());
            axpy(n, -H[i][j], V[i].data(), w.data());
        }
        H[j + 1][j] = norm2(n, w.data());
        if (H[j + 1][j] > EPS) {  // :128-131 (lucky breakdown leaves V[j+1]=0)
            V[j + 1] = w;
            scale(n, 1.0 / H[j + 1][j], V[j + 1].data());
        }
        for (int i = 0; i < j; ++i) rot(H[i][j], H[i + 1][j], cs[i], sn[i]);
        Real rv;
        givens(H[j][j], H[j + 1][j], cs[j], sn[j], rv);
        H[j][j] = rv;
        H[j + 1][j] = 0.0;
        rot(g[j], g[j + 1], cs[j], sn[j]);
        res = std::abs(g[j + 1]);
        if (res < tol_abs) { ++j; break; }
    }
    std::vector<Real> y(j, 0.0);  // solve_hessenberg, :57-66
    for (int i = j - 1; i >= 0; --i) {
        Real s = g[i];
        for (int k = i + 1; k < j; ++k) s -= H[i][k] * y[k];
        y[i] = std::abs(H[i][i]) > EPS ? s / H[i][i] : 0.0;
    }
    for (int i = 0; i < j; ++i) axpy(n, y[i], Z[i].data(), x);
    return j;
}

bool fgmres_solve(const Csr& A, const Real* b, Real* x, const Precond* M, const Params& prm, Stats& st) {
    const double t0 = now_s();
    const Index n = A.rows;
    const long long nz = A.nnz();
    const int m = std::min(prm.restart_size, n);
    const Real bn = norm2(n, b);
    if (bn < EPS) { st.converged = true; st.final_residual = 0.0; return true; }
    const Real tol_abs = prm.tol * bn;
    long long F = 0, B = 0;
    Real res = bn;
    int total = 0, restarts = 0;
    std::vector<std::vector<Real>> V, Z;
    while (total < prm.max_iter) {
        const int cm = std::min(m, prm.max_iter - total);
        const int its = fgmres_cycle(A, b, x, M, tol_abs, cm, res, V, Z);
        F += (long long)its * (spmv_flops(nz) + 4LL * n);
        B += (long long)its * (spmv_bytes(n, nz) + 4LL * n * 8);
        total += its;
        ++restarts;
        const Real rel = res / bn;
        if (rel < prm.tol || its == 0) { st.converged = rel < prm.tol; st.final_residual = rel; break; }
    }
    st.iterations = total;
    st.restarts = restarts - 1;
    st.solve_time_s = now_s() - t0;
    st.flop_count = F;
    st.mem_bytes = B;
    st.energy_joules = 2.0e-10 * F + 5.0e-9 * B;
    return st.converged;
}

// ---- spectral_interval: sparse_matrix.cpp:451-479 ---------------------------------
// K = min(n, steps) Lanczos steps from q = 1/sqrt(n); Gershgorin bounds of the
// tridiagonal T.
void spectral_interval(const Csr& A, int steps, Real* lo, Real* hi) {
    const Index n = A.rows;
    const int K = std::min<Index>(n, steps);
    std::vector<Real> q(n, 1.0 / std::sqrt((Real)n)), qp(n, 0.0), Aq(n);
    std::vector<Real> al, be;
    Real bj = 0.0;
    for (int j = 0; j < K; ++j) {
        spmv(A, q.data(), Aq.data());
        const Real aj = dot(n, q.data(), Aq.data());
        al.push_back(aj);
        for (Index i = 0; i < n; ++i) Aq[i] -= aj * q[i] + bj * qp[i];
        bj = norm2(n, Aq.data());
        if (bj < EPS * 1e6) break;
        be.push_back(bj);
        qp = q;
        q = Aq;
        scale(n, 1.0 / bj, q.data());
    }
    const int sz = (int)al.size();
    Real a = sz ? al[0] : 0.0, b = a;
    for (int i = 0; i < sz; ++i) {
        Real r = 0.0;
        if (i > 0) r += std::abs(be[i - 1]);
        if (i < sz - 1) r += std::abs(be[i]);
        a = std::min(a, al[i] - r);
        b = std::max(b, al[i] + r);
    }
    *lo = a;
    *hi = b;
}

// ---- s-step CG: sstep_cg_solver.cpp ----------------------------------------------------
// Each batch: Newton basis B_0 = r/|r|, B_j = (A - theta_{j-1}) B_{j-1}
// normalised (Chebyshev shifts on the Lanczos interval, :24-64); Gram G =
// B^T B (:69-78); if the diagonal ratio of G exceeds sstep_cond_tol, k is
// halved and the basis rebuilt (:217-227); G y = B^T r by Cholesky
// (:120-165), x += B y, r = b - A x exactly (:246-258).  A failed Cholesky
// takes one steepest-descent step with r instead (:232-243).
static std::vector<Real> cheb_shifts(Real lo, Real hi, int k) {  // :24-35
    std::vector<Real> s(k);
    const Real c = 0.5 * (lo + hi), h = 0.5 * (hi - lo);
    for (int j = 0; j < k; ++j) s[j] = c + h * std::cos(M_PI * (2.0 * j + 1.0) / (2.0 * k));
    return s;
}

static void sstep_basis(const Csr& A, const Real* r, const std::vector<Real>& sh,
                        std::vector<std::vector<Real>>& B) {  // :42-64
    const Index n = A.rows;
    const int k = (int)sh.size();
    B.assign(k, std::vector<Real>(n, 0.0));
    std::copy(r, r + n, B[0].begin());
    const Real n0 = norm2(n, B[0].data());
    if (n0 > EPS) scale(n, 1.0 / n0, B[0].data());
    std::vector<Real> tmp(n);
    for (int j = 1; j < k; ++j) {
        spmv(A, B[j - 1].data(), tmp.data());
        B[j] = tmp;
        axpy(n, -sh[j - 1], B[j - 1].data(), B[j].data());
        const Real nj = norm2(n, B[j].data());
        if (nj > EPS) scale(n, 1.0 / nj, B[j].data());
    }
}

static void sstep_gram(const std::vector<std::vector<Real>>& B, std::vector<std::vector<Real>>& G) {  // :69-78
    const int k = (int)B.size();
    const Index n = k ? (Index)B[0].size() : 0;
    G.assign(k, std::vector<Real>(k, 0.0));
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j) G[i][j] = G[j][i] = dot(n, B[i].data(), B[j].data());
}

static double sstep_cond(const std::vector<std::vector<Real>>& G) {  // :103-114
    Real mx = 0.0, mn = std::numeric_limits<Real>::infinity();
    for (size_t i = 0; i < G.size(); ++i) {
        mx = std::max(mx, std::abs(G[i][i]));
        if (std::abs(G[i][i]) > EPS) mn = std::min(mn, std::abs(G[i][i]));
    }
    return mn >= mx ? 1.0 : mx / mn;
}

static bool sstep_cholesky(const std::vector<std::vector<Real>>& G, const std::vector<Real>& rhs,
                           std::vector<Real>& y) {  // :120-165
    const int k = (int)G.size();
    std::vector<std::vector<Real>> L(k, std::vector<Real>(k, 0.0));
    for (int i = 0; i < k; ++i)
        for (int j = 0; j <= i; ++j) {
            Real s = G[i][j];
            for (int l = 0; l < j; ++l) s -= L[i][l] * L[j][l];
            if (i == j) {
                if (s <= 0.0) return false;
                L[i][j] = std::sqrt(s);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
    std::vector<Real> z(k, 0.0);
    for (int i = 0; i < k; ++i) {
        Real s = rhs[i];
        for (int j = 0; j < i; ++j) s -= L[i][j] * z[j];
        z[i] = s / L[i][i];
    }
    y.assign(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        Real s = z[i];
        for (int j = i + 1; j < k; ++j) s -= L[j][i] * y[j];
        y[i] = s / L[i][i];
    }
    return true;
}

bool sstep_cg_solve(const Csr& A, const Real* b, Real* x, const Params& prm, Stats& st) {  // :168-305
    const double t0 = now_s();
    const Index n = A.rows;
    const long long nz = A.nnz();
    const int k = std::min(prm.sstep_k, std::max(1, n / 10));
    Real lo_raw, hi_raw;
    spectral_interval(A, 20, &lo_raw, &hi_raw);
    const Real lo = std::max(lo_raw, EPS);
    const Real hi = std::max(hi_raw, lo * 2.0);
    std::vector<Real> r(n);
    spmv(A, x, r.data());
    for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const Real bn = norm2(n, b);
    Real rn = norm2(n, r.data());
    st.initial_residual = rn;
    if (bn < EPS) { st.converged = true; st.final_residual = 0.0; return true; }
    Real rel = rn / bn;
    if (rel < prm.tol) { st.converged = true; st.final_residual = rel; return true; }
    long long F = 0, Bb = 0;
    int total = 0;
    std::vector<std::vector<Real>> B, G;
    while (total < prm.max_iter && rel > prm.tol) {
        int kb = k;
        sstep_basis(A, r.data(), cheb_shifts(lo, hi, kb), B);
        F += (long long)kb * spmv_flops(nz);
        Bb += (long long)kb * spmv_bytes(n, nz);
        sstep_gram(B, G);
        if (sstep_cond(G) > prm.sstep_cond_tol && kb > 1) {
            kb = std::max(1, kb / 2);
            sstep_basis(A, r.data(), cheb_shifts(lo, hi, kb), B);
            sstep_gram(B, G);
        }
        std::vector<Real> rhs(kb), y;
        for (int i = 0; i < kb; ++i) rhs[i] = dot(n, r.data(), B[i].data());
        F += (long long)kb * 2LL * n;
        if (!sstep_cholesky(G, rhs, y)) {
            std::vector<Real> Ap(n);
            spmv(A, r.data(), Ap.data());
            const Real rr = dot(n, r.data(), r.data()), rAp = dot(n, r.data(), Ap.data());
            if (std::abs(rAp) > EPS) {
                const Real alpha = rr / rAp;
                axpy(n, alpha, r.data(), x);
                axpy(n, -alpha, Ap.data(), r.data());
            }
            ++total;
            rel = norm2(n, r.data()) / bn;
            continue;
        }
        for (int i = 0; i < kb; ++i) axpy(n, y[i], B[i].data(), x);
        spmv(A, x, r.data());
        F += spmv_flops(nz);
        Bb += spmv_bytes(n, nz);
        for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
        rn = norm2(n, r.data());
        rel = rn / bn;
        total += kb;
    }
    st.iterations = total;
    st.final_residual = rel;
    st.converged = rel < prm.tol;
    st.solve_time_s = now_s() - t0;
    st.flop_count = F;
    st.mem_bytes = Bb;
    st.energy_joules = 2e-10 * F + 5e-9 * Bb;
    return st.converged;
}

// ---- pipelined FGMRES: pipelined_fgmres_solver.cpp:56-243 ---------------------------
// The FGMRES cycle with the next direction's preconditioner apply and SpMV
// issued right after it is normalised (the same values), a lucky breakdown
// ending the cycle before column j is rotated (:165-168), convergence on the
// Givens residual, and (its + 1) SpMVs charged per cycle (:224-227).
static int pipelined_cycle(const Csr& A, const Real* b, Real* x, const Precond* M, Real tol_abs, int m, Real& res) {
    const Index n = A.rows;
    std::vector<Real> r(n);
    spmv(A, x, r.data());
    for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const Real beta = norm2(n, r.data());
    res = beta;
    if (beta < tol_abs) return 0;
    std::vector<std::vector<Real>> V(m + 1, std::vector<Real>(n, 0.0)), Z(m, std::vector<Real>(n, 0.0)),
        S(m, std::vector<Real>(n, 0.0));
    V[0] = r;
    scale(n, 1.0 / beta, V[0].data());
    if (M) M->apply(V[0].data(), Z[0].data()); else Z[0] = V[0];
    spmv(A, Z[0].data(), S[0].data());
    std::vector<std::vector<Real>> H(m + 1, std::vector<Real>(m, 0.0));
    std::vector<Real> cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m; ++j) {
        std::vector<Real>& w = S[j];
        for (int i = 0; i <= j; ++i) {
            H[i][j] = dot(n, w.data(), V[i].data());
            axpy(n, -H[i][j], V[i].data(), w.data());
        }
        H[j + 1][j] = norm2(n, w.data());
        if (H[j + 1][j] > EPS) {
            V[j + 1] = w;
            scale(n, 1.0 / H[j + 1][j], V[j + 1].data());
        } else {
            ++j;
            break;
        }
        if (j + 1 < m) {
            if (M) M->apply(V[j + 1].data(), Z[j + 1].data()); else Z[j + 1] = V[j + 1];
            spmv(A, Z[j + 1].data(), S[j + 1].data());
        }
        for (int i = 0; i < j; ++i) rot(H[i][j], H[i + 1][j], cs[i], sn[i]);
        Real rv;
        givens(H[j][j], H[j + 1][j], cs[j], sn[j], rv);
        H[j][j] = rv;
        H[j + 1][j] = 0.0;
        rot(g[j], g[j + 1], cs[j], sn[j]);
        res = std::abs(g[j + 1]);
        if (res < tol_abs) { ++j; break; }
    }
    std::vector<Real> y(j, 0.0);
    for (int i = j - 1; i >= 0; --i) {
        Real s = g[i];
        for (int q = i + 1; q < j; ++q) s -= H[i][q] * y[q];
        y[i] = std::abs(H[i][i]) > EPS ? s / H[i][i] : 0.0;
    }
    for (int i = 0; i < j; ++i) axpy(n, y[i], Z[i].data(), x);
    return j;
}

bool pipelined_fgmres_solve(const Csr& A, const Real* b, Real* x, const Precond* M, const Params& prm, Stats& st) {
    const double t0 = now_s();
    const Index n = A.rows;
    const long long nz = A.nnz();
    const int m = std::min(prm.restart_size, n);
    const Real bn = norm2(n, b);
    if (bn < EPS) { st.converged = true; st.final_residual = 0.0; return true; }
    const Real tol_abs = prm.tol * bn;
    long long F = 0, B = 0;
    Real res = bn;
    int total = 0, restarts = 0;
    while (total < prm.max_iter) {
        const int cm = std::min(m, prm.max_iter - total);
        const int its = pipelined_cycle(A, b, x, M, tol_abs, cm, res);
        F += (long long)(its + 1) * spmv_flops(nz) + (long long)its * 4LL * n;
        B += (long long)(its + 1) * spmv_bytes(n, nz) + (long long)its * 4LL * n * 8;
        total += its;
        ++restarts;
        const Real rel = res / bn;
        if (rel < prm.tol || its == 0) { st.converged = rel < prm.tol; st.final_residual = rel; break; }
    }
    st.iterations = total;
    st.restarts = restarts - 1;
    st.solve_time_s = now_s() - t0;
    st.flop_count = F;
    st.mem_bytes = B;
    st.energy_joules = 2.0e-10 * F + 5.0e-9 * B;
    return st.converged;
}

// ---- factory: solver_factory.cpp:55-133 ----------------------------------------------
std::unique_ptr<Precond> make_precond(int kind, const Csr& A, const Params& p) {
    switch (kind) {
        case 0: return nullptr;
        case 1: { auto j = std::make_unique<Jacobi>(); j->setup(A); return j; }
        case 2: { auto i = std::make_unique<Ilu>(); i->drop_tol = p.drop_tol; i->fill_p = 0; i->setup(A); return i; }
        case 3: {
            auto a = std::make_unique<Amg>();
            a->max_levels = p.amg_levels;
            a->theta = p.amg_strong;
            a->pre = p.amg_smooth_pre;
            a->post = p.amg_smooth_post;
            a->setup(A);
            return a;
        }
    }
    throw std::invalid_argument("SolverFactory: unknown PrecondType");
}

// ---- extract_features: adaptive_selector.cpp:38-220 (fields the ladder reads) ---------
Features extract_features(const Csr& A) {
    Features f;
    f.n = A.rows;
    f.nnz = A.nnz();
    f.density = (A.rows == 0 || A.cols == 0) ? 0.0 : (double)A.nnz() / ((double)A.rows * A.cols);
    Real fro = 0.0;
    for (Real v : A.v) fro += v * v;
    f.frobenius = std::sqrt(fro);
    Real min_dd = std::numeric_limits<Real>::infinity();
    for (Index i = 0; i < f.n; ++i) {
        Real d = 0.0, off = 0.0;
        for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            if (A.ci[k] == i) d = std::abs(A.v[k]); else off += std::abs(A.v[k]);
        }
        const Real dd = off > EPS ? d / off : std::numeric_limits<Real>::infinity();
        if (dd < min_dd) min_dd = dd;
    }
    f.diag_dominance = std::isinf(min_dd) ? 1.0 : min_dd;
    {  // 30-step Lanczos + Gershgorin on T, :66-126
        const int K = std::min(f.n, 30);
        const Index n = f.n;
        std::vector<Real> q(n, 1.0 / std::sqrt((Real)n)), qp(n, 0.0), Aq(n);
        std::vector<Real> al, be;
        Real bj = 0.0;
        for (int j = 0; j < K; ++j) {
            spmv(A, q.data(), Aq.data());
            const Real aj = dot(n, q.data(), Aq.data());
            al.push_back(aj);
            for (Index i = 0; i < n; ++i) Aq[i] -= aj * q[i] + bj * qp[i];
            bj = norm2(n, Aq.data());
            if (bj < EPS * 1e6) break;
            be.push_back(bj);
            qp = q;
            q = Aq;
            scale(n, 1.0 / bj, q.data());
        }
        const int sz = (int)al.size();
        Real tmin = al[0], tmax = al[0];
        for (int i = 0; i < sz; ++i) {
            Real r = 0.0;
            if (i > 0) r += std::abs(be[i - 1]);
            if (i < sz - 1) r += std::abs(be[i]);
            tmin = std::min(tmin, al[i] - r);
            tmax = std::max(tmax, al[i] + r);
        }
        const Real lmax = std::max(tmax, EPS);
        Real lmin = tmin;
        if (lmin <= EPS) {
            Real mpd = std::numeric_limits<Real>::infinity();
            for (Index i = 0; i < n; ++i)
                for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
                    if (A.ci[k] == i && A.v[k] > EPS) { mpd = std::min(mpd, A.v[k]); break; }
            if (std::isinf(mpd)) mpd = EPS;
            lmin = mpd * 0.5;
        }
        lmin = std::max(lmin, EPS);
        f.estimated_cond = lmax / lmin;
        f.lambda_max = lmax;
        f.lambda_min = lmin;
    }
    {  // off_diag_symmetry over the first 200 rows, :167-178
        Real mx = 0.0;
        for (Index i = 0; i < std::min(f.n, 200); ++i)
            for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const Index j = A.ci[k];
                if (j == i) continue;
                mx = std::max(mx, std::abs(A.v[k] - get(A, j, i)));
            }
        f.symmetry_residual = f.frobenius > EPS ? mx / f.frobenius : 0.0;
    }
    {  // SPD sample, :200-219
        bool sym = true;
        const Index sr = std::min((Index)200, f.n);
        const Index stride = std::max((Index)1, f.n / sr);
        for (Index i = 0; i < f.n && sym; i += stride)
            for (Index k = A.rp[i]; k < A.rp[i + 1] && sym; ++k) {
                const Index j = A.ci[k];
                if (j == i) continue;
                if (std::abs(A.v[k] - get(A, j, i)) > 1e-6 * (1.0 + std::abs(A.v[k]))) sym = false;
            }
        bool pd = true;
        for (Index i = 0; i < f.n && pd; ++i)
            for (Index k = A.rp[i]; k < A.rp[i + 1]; ++k)
                if (A.ci[k] == i) { if (A.v[k] <= 0.0) pd = false; break; }
        f.is_spd = sym && pd;
    }
    return f;
}

int select_initial_state(const Features& f) {  // adaptive_selector.cpp:225-233
    if (f.is_spd && f.estimated_cond < 1e6 && f.diag_dominance >= 0.8) return 0;
    if (f.estimated_cond >= 1e6 || f.diag_dominance < 0.3) return 2;
    return 1;
}

int feature_advise(const Features& f, Params& p) {  // ml_advisor.cpp:38-100, ml_advisor.hpp:109-113
    const bool sym = f.symmetry_residual < 1e-4;
    int tier;
    if (f.estimated_cond < 1e4 && f.diag_dominance >= 0.8 && sym && f.is_spd) tier = 0;
    else if (f.estimated_cond >= 1e7 || f.diag_dominance < 0.3 || !sym) tier = 2;
    else tier = 1;
    if (tier == 2) p.restart_size = std::min(p.restart_size * 2, f.n);
    else if (tier == 1) p.restart_size = std::min((int)(p.restart_size * 1.5), f.n);
    if (f.estimated_cond > 1e5) p.drop_tol = 1e-5;
    else if (f.estimated_cond > 1e3) p.drop_tol = 1e-4;
    else p.drop_tol = 1e-3;
    if (f.diag_dominance > 0.9) p.amg_strong = 0.5;
    else if (f.diag_dominance > 0.5) p.amg_strong = 0.25;
    else p.amg_strong = 0.1;
    p.amg_levels = std::max(5, (int)std::log2((double)f.n / 50.0) + 2);
    return tier;
}

// ---- AdaptiveSelector::solve: adaptive_selector.cpp:298-394 --------------------------
bool adaptive_solve(const Csr& A, const Real* b, Real* x, Params p, int advisor, Stats& st) {
    Features f = extract_features(A);
    int state = advisor == 1 ? feature_advise(f, p) : select_initial_state(f);
    bool solved = false;
    int esc = 0;
    while (!solved) {
        const double t0 = now_s();
        auto M = make_precond(state == 0 ? 1 : state == 1 ? 2 : 3, A, p);
        const double setup = now_s() - t0;
        Stats at;
        at.setup_time_s = setup;
        solved = state == 0 ? cg_solve(A, b, x, M.get(), p, at) : fgmres_solve(A, b, x, M.get(), p, at);
        st = at;
        st.setup_time_s += at.setup_time_s;  // doubled, :350-351
        if (solved || state == 2) break;
        ++state;
        ++esc;
    }
    st.state = state;
    st.escalations = esc;
    return solved;
}

}  // namespace orc
