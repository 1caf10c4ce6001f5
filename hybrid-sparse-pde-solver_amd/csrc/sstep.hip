// =============================================================================
// sstep.hip — s-step CG (sstep_cg_solver.cpp:168-305) and the Lanczos
// spectral interval it starts from (sparse_matrix.cpp:451-479), on the device.
//
// Per batch of k steps:
//   * the Newton basis B_0 = r/|r|, B_j = (A - theta_{j-1}) B_{j-1} / |.|
//     (Chebyshev shifts on the Lanczos interval, :24-64): one SpMV per vector
//     with the shift and the |B_j|^2 partials in its epilogue, a finish and a
//     scale;
//   * the Gram matrix B^T B and the right-hand side B^T r in ONE fused pass
//     over the basis (:69-78, :229-231): up to 16 dot products per launch;
//   * one readback per batch; the host takes the diagonal-ratio condition
//     estimate (:103-114, halving k and rebuilding above sstep_cond_tol),
//     solves the k x k system by Cholesky (:120-165), or, when it fails,
//     takes the one steepest-descent step of :232-243;
//   * x += B y in one pass, then r = b - A x with |r|^2 (one more readback).
// The reference's point — one global reduction per k steps instead of one
// per step — becomes one host round trip per k steps here.  As in the
// reference, the preconditioner is not used.
// =============================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "device_common.hpp"

namespace hsps_dev {
namespace {

constexpr double kEpsS = std::numeric_limits<double>::epsilon();
constexpr int kPairs = 16;  // dot products / axpys per fused launch
constexpr int kLanczosMax = 64;

struct Pairs {
    const double* a[kPairs];
    const double* b[kPairs];
    int np;
};
struct Axpys {
    const double* v[kPairs];
    double c[kPairs];
    int nv;
};

// Partials of a_k . b_k for k < np (np <= 16) in one pass.
__global__ void __launch_bounds__(kBlock) k_multi_dot(int n, Pairs P, const int* skip, double* partials) {
    if (skip && *skip) return;
    double red[kPairs];
#pragma unroll
    for (int k = 0; k < kPairs; ++k) red[k] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < kPairs; ++k)
            if (k < P.np) red[k] += P.a[k][i] * P.b[k][i];
    }
    block_partial<kPairs>(red, partials);
}

// x += c_0 v_0; x += c_1 v_1; ... element by element, in order (x.axpy per
// basis vector, :250-251).
__global__ void k_multi_axpy(int n, double* __restrict__ x, Axpys A) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = x[i];
#pragma unroll
        for (int k = 0; k < kPairs; ++k)
            if (k < A.nv) v += A.c[k] * A.v[k][i];
        x[i] = v;
    }
}

// y = A x - theta x (B_j = A B_{j-1}, then B_j += -theta B_{j-1}, :54-57);
// |y|^2 partials.
struct EpiShift : SkipFlag {
    static constexpr int NV = 1;
    const double* x;
    double* y;
    double neg_theta;
    double* partials = nullptr;
    template <class A> __device__ void row(int i, A s, double* red) const {
        const double v = (double)s + neg_theta * x[i];
        y[i] = v;
        red[0] += v * v;
    }
};

// y = x * (1 / sqrt(*nrm2)) when sqrt(*nrm2) > eps, else y = x (:48-49, 60-61).
__global__ void k_scale_inv_norm(int n, const double* __restrict__ x, double* __restrict__ y, const double* nrm2) {
    const double nr = sqrt(*nrm2);
    const bool sc = nr > kEpsS;
    const double f = sc ? 1.0 / nr : 1.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = sc ? x[i] * f : x[i];
}

__global__ void k_fill(int n, double* __restrict__ y, double v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = v;
}

// ---- Lanczos (sparse_matrix.cpp:451-479) -------------------------------------------
struct LzState {
    int stop;
    int na;  // alphas recorded
    double alpha[kLanczosMax];
    double beta[kLanczosMax];
};

struct OpLzAlpha {
    LzState* st;
    int j;
    __device__ bool skip() const { return st->stop != 0; }
    __device__ void operator()(const double* tot) const {
        st->alpha[j] = tot[0];
        st->na = j + 1;
    }
};

// Aq -= alpha_j q + beta_{j-1} q_prev; |Aq|^2 partials
__global__ void __launch_bounds__(kBlock) k_lz_update(int n, double* __restrict__ aq, const double* __restrict__ q,
                                                      const double* __restrict__ qp, const LzState* st, int j,
                                                      double* partials) {
    if (st->stop) return;
    const double a = st->alpha[j], b = j > 0 ? st->beta[j - 1] : 0.0;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double v = aq[i] - (a * q[i] + b * qp[i]);
        aq[i] = v;
        red[0] += v * v;
    }
    block_partial<1>(red, partials);
}

struct OpLzBeta {
    LzState* st;
    int j;
    __device__ bool skip() const { return st->stop != 0; }
    __device__ void operator()(const double* tot) const {
        const double bj = sqrt(tot[0]);
        if (bj < kEpsS * 1e6) {
            st->stop = 1;
            return;
        }
        st->beta[j] = bj;
    }
};

// q_prev = q; q = Aq * (1 / beta_j)
__global__ void k_lz_next(int n, double* __restrict__ q, double* __restrict__ qp, const double* __restrict__ aq,
                          const LzState* st, int j) {
    if (st->stop) return;
    const double f = 1.0 / st->beta[j];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        qp[i] = q[i];
        q[i] = aq[i] * f;
    }
}

double* dptr(DevBuf& b, size_t off = 0) { return b.as<double>() + off; }

void read_back(void* h, const void* d, size_t bytes, hipStream_t s) {
    HSPS_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    HSPS_HIP(hipStreamSynchronize(s));
}

// Dot products of the given pairs into out[0..np) (device), 16 per launch.
void multi_dot(int n, const std::vector<std::pair<const double*, const double*>>& pr, double* out, double* part,
               int g, hipStream_t s) {
    for (size_t b = 0; b < pr.size(); b += kPairs) {
        Pairs P{};
        P.np = (int)std::min<size_t>(kPairs, pr.size() - b);
        for (int k = 0; k < P.np; ++k) {
            P.a[k] = pr[b + k].first;
            P.b[k] = pr[b + k].second;
        }
        hipLaunchKernelGGL(k_multi_dot, dim3(g), dim3(kBlock), 0, s, n, P, nullptr, part);
        launch_finish<kPairs>(part, g, OpStore<kPairs>{out + b}, s);
        HSPS_CHECK_LAUNCH();
    }
}

}  // namespace

void spectral_interval_dev(const Csr& A, int steps, double* lo, double* hi) {
    Ctx* c = A.ctx;
    hipStream_t s = c->stream;
    const int n = A.rows;
    if (n == 0) {
        *lo = *hi = 0.0;
        return;
    }
    const int K = std::min(std::min(n, steps), kLanczosMax);
    const int g = reduction_blocks(n);
    DevBuf q(c, (size_t)n * 8), qp(c, (size_t)n * 8), aq(c, (size_t)n * 8);
    DevBuf part(c, sizeof(double) * (size_t)std::max(std::max(g, A.n_blocks), 1) * kPairs);
    DevBuf stb(c, sizeof(LzState));
    HSPS_HIP(hipMemsetAsync(stb.p, 0, sizeof(LzState), s));
    LzState* st = stb.as<LzState>();
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dptr(q), 1.0 / std::sqrt((double)n));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dptr(qp), 0.0);
    HSPS_CHECK_LAUNCH();
    for (int j = 0; j < K; ++j) {
        spmv_f64(A, dptr(q), dptr(aq), s, &st->stop);
        Pairs P{};
        P.np = 1;
        P.a[0] = dptr(q);
        P.b[0] = dptr(aq);
        hipLaunchKernelGGL(k_multi_dot, dim3(g), dim3(kBlock), 0, s, n, P, &st->stop, dptr(part));
        launch_finish<kPairs>(dptr(part), g, OpLzAlpha{st, j}, s);
        hipLaunchKernelGGL(k_lz_update, dim3(g), dim3(kBlock), 0, s, n, dptr(aq), dptr(q), dptr(qp), st, j, dptr(part));
        launch_finish<1>(dptr(part), g, OpLzBeta{st, j}, s);
        hipLaunchKernelGGL(k_lz_next, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dptr(q), dptr(qp), dptr(aq), st, j);
        HSPS_CHECK_LAUNCH();
    }
    LzState h{};
    read_back(&h, st, sizeof h, s);
    const int sz = h.na;
    double a = sz ? h.alpha[0] : 0.0, b = a;
    for (int i = 0; i < sz; ++i) {  // Gershgorin discs of T
        double r = 0.0;
        if (i > 0) r += std::abs(h.beta[i - 1]);
        if (i < sz - 1) r += std::abs(h.beta[i]);
        a = std::min(a, h.alpha[i] - r);
        b = std::max(b, h.alpha[i] + r);
    }
    *lo = a;
    *hi = b;
}

int sstep_cg_dev(const Csr& A, const double* b, double* x, const hsps_params& p, hsps_stats* out) {
    Ctx* c = A.ctx;
    hipStream_t s = c->stream;
    const int n = A.rows;
    const long long nz = A.nnz;
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    auto spmv_bytes = [&](long long rows) { return (rows + 1 + nz) * 4 + nz * 8 + rows * 16; };
    hsps_stats st{};
    const int k = std::max(1, std::min(p.sstep_k, std::max(1, n / 10)));
    st.sstep_k = k;
    double lo_raw = 0.0, hi_raw = 0.0;
    spectral_interval_dev(A, 20, &lo_raw, &hi_raw);
    const double lo = std::max(lo_raw, kEpsS);
    const double hi = std::max(hi_raw, lo * 2.0);
    const size_t ld = ((size_t)std::max(n, 1) + 63) / 64 * 64;
    const int g = reduction_blocks(n);
    DevBuf basis(c, ld * (size_t)k * 8), r(c, ld * 8), ap(c, ld * 8);
    DevBuf part(c, sizeof(double) * (size_t)std::max(std::max(g, A.n_blocks), 1) * kPairs);
    const int npairs_max = k * (k + 1) / 2 + k;
    DevBuf dots(c, sizeof(double) * (size_t)(npairs_max + kPairs));
    double* D = dots.as<double>();
    double* P0 = part.as<double>();
    auto Bv = [&](int i) { return basis.as<double>() + (size_t)i * ld; };
    auto finish = [&](int conv) {
        st.converged = conv;
        st.solve_time_s = elapsed();
        if (out) *out = st;
        return conv;
    };
    // r = b - A x with |r|^2; |b|^2
    residual_f64(A, b, x, r.as<double>(), s, D, P0, nullptr);
    multi_dot(n, {{b, b}}, D + 1, P0, g, s);
    double h2[2];
    read_back(h2, D, sizeof h2, s);
    const double bn = std::sqrt(h2[1]);
    double rn = std::sqrt(h2[0]);
    st.initial_residual = rn;
    if (bn < kEpsS) {
        st.final_residual = 0.0;
        st.true_residual = 0.0;
        return finish(1);
    }
    double rel = rn / bn;
    if (rel < p.tol) {
        st.final_residual = rel;
        st.true_residual = rel;
        return finish(1);
    }
    long long F = 0, Bytes = 0;
    int total = 0, batches = 0;
    auto shifts_for = [&](int kb) {
        std::vector<double> sh(kb);
        const double cc = 0.5 * (lo + hi), hh = 0.5 * (hi - lo);
        for (int j = 0; j < kb; ++j) sh[j] = cc + hh * std::cos(M_PI * (2.0 * j + 1.0) / (2.0 * kb));
        return sh;
    };
    auto build_basis = [&](int kb) {
        const std::vector<double> sh = shifts_for(kb);
        multi_dot(n, {{r.as<double>(), r.as<double>()}}, D, P0, g, s);
        hipLaunchKernelGGL(k_scale_inv_norm, dim3(grid_for(n)), dim3(kBlock), 0, s, n, (const double*)r.as<double>(),
                           Bv(0), (const double*)D);
        HSPS_CHECK_LAUNCH();
        for (int j = 1; j < kb; ++j) {
            EpiShift e;
            e.x = Bv(j - 1);
            e.y = Bv(j);
            e.neg_theta = -sh[j - 1];
            e.partials = P0;
            launch_spmv(A, A.v64.as<double>(), (const double*)Bv(j - 1), e, s);
            launch_finish<1>(P0, A.n_blocks, OpStore<1>{D}, s);
            hipLaunchKernelGGL(k_scale_inv_norm, dim3(grid_for(n)), dim3(kBlock), 0, s, n, (const double*)Bv(j), Bv(j),
                               (const double*)D);
            HSPS_CHECK_LAUNCH();
        }
    };
    // Gram (upper triangle, row-major pairs) and B^T r in one fused sweep
    auto gram = [&](int kb, std::vector<std::vector<double>>& G, std::vector<double>& rhs) {
        std::vector<std::pair<const double*, const double*>> pr;
        for (int i = 0; i < kb; ++i)
            for (int j = i; j < kb; ++j) pr.push_back({Bv(i), Bv(j)});
        for (int i = 0; i < kb; ++i) pr.push_back({r.as<double>(), Bv(i)});
        multi_dot(n, pr, D, P0, g, s);
        std::vector<double> h(pr.size());
        read_back(h.data(), D, h.size() * sizeof(double), s);
        G.assign(kb, std::vector<double>(kb, 0.0));
        size_t q = 0;
        for (int i = 0; i < kb; ++i)
            for (int j = i; j < kb; ++j, ++q) G[i][j] = G[j][i] = h[q];
        rhs.assign(h.begin() + q, h.end());
    };
    while (total < p.max_iter && rel > p.tol) {
        int kb = k;
        build_basis(kb);
        F += (long long)kb * 2LL * nz;
        Bytes += (long long)kb * spmv_bytes(n);
        std::vector<std::vector<double>> G;
        std::vector<double> rhs, y;
        gram(kb, G, rhs);
        double mx = 0.0, mn = std::numeric_limits<double>::infinity();  // :103-114
        for (int i = 0; i < kb; ++i) {
            mx = std::max(mx, std::abs(G[i][i]));
            if (std::abs(G[i][i]) > kEpsS) mn = std::min(mn, std::abs(G[i][i]));
        }
        const double cond = mn >= mx ? 1.0 : mx / mn;
        if (cond > p.sstep_cond_tol && kb > 1) {
            kb = std::max(1, kb / 2);
            build_basis(kb);
            gram(kb, G, rhs);
        }
        ++batches;
        F += (long long)kb * 2LL * n;
        // Cholesky G = L L^T, then L z = rhs, L^T y = z (:120-165)
        bool ok = true;
        std::vector<std::vector<double>> L(kb, std::vector<double>(kb, 0.0));
        for (int i = 0; i < kb && ok; ++i)
            for (int j = 0; j <= i; ++j) {
                double sum = G[i][j];
                for (int l = 0; l < j; ++l) sum -= L[i][l] * L[j][l];
                if (i == j) {
                    if (sum <= 0.0) { ok = false; break; }
                    L[i][j] = std::sqrt(sum);
                } else {
                    L[i][j] = sum / L[j][j];
                }
            }
        if (ok) {
            std::vector<double> z(kb, 0.0);
            for (int i = 0; i < kb; ++i) {
                double sum = rhs[i];
                for (int j = 0; j < i; ++j) sum -= L[i][j] * z[j];
                z[i] = sum / L[i][i];
            }
            y.assign(kb, 0.0);
            for (int i = kb - 1; i >= 0; --i) {
                double sum = z[i];
                for (int j = i + 1; j < kb; ++j) sum -= L[j][i] * y[j];
                y[i] = sum / L[i][i];
            }
        } else {
            // one steepest-descent step with r (:232-243)
            spmv_f64(A, r.as<double>(), ap.as<double>(), s);
            multi_dot(n, {{r.as<double>(), r.as<double>()}, {r.as<double>(), ap.as<double>()}}, D, P0, g, s);
            double rr[2];
            read_back(rr, D, sizeof rr, s);
            if (std::abs(rr[1]) > kEpsS) {
                const double alpha = rr[0] / rr[1];
                Axpys X{};
                X.nv = 1;
                X.v[0] = r.as<double>();
                X.c[0] = alpha;
                hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, X);
                X.v[0] = ap.as<double>();
                X.c[0] = -alpha;
                hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r.as<double>(), X);
                HSPS_CHECK_LAUNCH();
            }
            ++total;
            multi_dot(n, {{r.as<double>(), r.as<double>()}}, D, P0, g, s);
            double r2 = 0.0;
            read_back(&r2, D, sizeof r2, s);
            rel = std::sqrt(r2) / bn;
            continue;
        }
        for (int b0 = 0; b0 < kb; b0 += kPairs) {
            Axpys X{};
            X.nv = std::min(kPairs, kb - b0);
            for (int q = 0; q < X.nv; ++q) {
                X.v[q] = Bv(b0 + q);
                X.c[q] = y[b0 + q];
            }
            hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, X);
            HSPS_CHECK_LAUNCH();
        }
        residual_f64(A, b, x, r.as<double>(), s, D, P0, nullptr);
        F += 2LL * nz;
        Bytes += spmv_bytes(n);
        double r2 = 0.0;
        read_back(&r2, D, sizeof r2, s);
        rn = std::sqrt(r2);
        rel = rn / bn;
        total += kb;
    }
    st.iterations = total;
    st.restarts = 0;
    st.comm_events = batches;
    st.final_residual = rel;
    residual_f64(A, b, x, ap.as<double>(), s, D, P0, nullptr);
    double t2 = 0.0;
    read_back(&t2, D, sizeof t2, s);
    st.true_residual = std::sqrt(t2) / bn;
    st.flop_count = F;
    st.mem_bytes = Bytes;
    st.energy_joules = 2e-10 * (double)F + 5e-9 * (double)Bytes;
    return finish(rel < p.tol ? 1 : 0);
}

}  // namespace hsps_dev
