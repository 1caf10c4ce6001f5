// =============================================================================
// precond.hip — device preconditioners behind PreconditionerBase
// (preconditioner_base.hpp:16-48):
//   Jacobi   z = omega D^-1 r                      (jacobi.cpp:13-51)
//   ILU      L\U from ILU(0)/ILUT (host, exact), level-scheduled device
//            forward/backward sweeps               (ilu.cpp:31-271)
//   SA-AMG   host aggregation (bit-exact), smoothed P, R = P^T, Galerkin;
//            device V(pre,post) cycle with fused smoother/residual SpMV
//            epilogues and a dense coarsest operator (amg.cpp:30-343)
// =============================================================================
#include <chrono>
#include <cmath>
#include <limits>
#include <type_traits>

#include "dense_lu.hpp"
#include "device_common.hpp"
#include "host_setup.hpp"
#include "precond.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

// ---- element-wise kernels ---------------------------------------------------------
template <class T>
__global__ void k_diag_scale(int n, const T* __restrict__ d, const T* __restrict__ r, T* __restrict__ z,
                             const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) z[i] = d[i] * r[i];
}

// x = omega * dinv * rhs : the first smoothing sweep from a zero guess
// (A*0 = 0 in amg.cpp:247-252).
template <class T>
__global__ void k_first_sweep(int n, T omega, const T* __restrict__ dinv, const T* __restrict__ rhs,
                              T* __restrict__ x, const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = omega * dinv[i] * (rhs[i] - T(0));
}

template <class From, class To>
__global__ void k_convert(int n, const From* __restrict__ a, To* __restrict__ b, const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (To)a[i];
}

template <class T>
__global__ void k_copy(int n, const T* __restrict__ a, T* __restrict__ b, const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

// Dense z = G rhs (FP64 G, row-major n x n with row stride ld), one wave64
// per row, FP64 arithmetic in both precisions (the coarse solve of the mixed
// V-cycle stays FP64).
template <class T>
__global__ void __launch_bounds__(kBlock)
k_dense_gemv(int n, int ld, const double* __restrict__ G, const T* __restrict__ rhs, T* __restrict__ z,
             const int* skip) {
    if (skip && *skip) return;
    const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    const double* g = G + (size_t)row * ld;
    for (int c = lane; c < n; c += 64) s += (double)g[c] * (double)rhs[c];
    s = wave_sum(s);
    if (lane == 0) z[row] = (T)s;
}

// Level-scheduled triangular sweeps in one workgroup: rows of a level are
// independent; a workgroup barrier separates levels.  Forward: unit lower
// (ilu.cpp:225-239, columns sorted so the scan stops at j >= i).  Backward:
// divide by U(i,i) (ilu.cpp:244-262).
constexpr int kTriThreads = 1024;
template <class V, class T>
__global__ void __launch_bounds__(kTriThreads)
k_tri_lower(int n_levels, const int* __restrict__ lptr, const int* __restrict__ order, const int* __restrict__ rp,
            const int* __restrict__ ci, const V* __restrict__ lu, const T* __restrict__ r, T* y, const int* skip) {
    if (skip && *skip) return;
    for (int lev = 0; lev < n_levels; ++lev) {
        const int a = lptr[lev], b = lptr[lev + 1];
        for (int idx = a + (int)threadIdx.x; idx < b; idx += kTriThreads) {
            const int i = order[idx];
            T s = r[i];
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int j = ci[k];
                if (j >= i) break;
                s -= (T)lu[k] * y[j];
            }
            y[i] = s;
        }
        __syncthreads();
    }
}
template <class V, class T>
__global__ void __launch_bounds__(kTriThreads)
k_tri_upper(int n_levels, const int* __restrict__ lptr, const int* __restrict__ order, const int* __restrict__ rp,
            const int* __restrict__ ci, const V* __restrict__ lu, const T* __restrict__ inv_d, const T* __restrict__ y,
            T* z, const int* skip) {
    if (skip && *skip) return;
    for (int lev = 0; lev < n_levels; ++lev) {
        const int a = lptr[lev], b = lptr[lev + 1];
        for (int idx = a + (int)threadIdx.x; idx < b; idx += kTriThreads) {
            const int i = order[idx];
            T s = y[i];
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int j = ci[k];
                if (j > i) s -= (T)lu[k] * z[j];
            }
            z[i] = s * inv_d[i];
        }
        __syncthreads();
    }
}

// The same sweeps over the level-ordered ELL structure (TriEll): each row's
// dependencies and values load independently of one another, a dependency on
// the previous level reads the LDS copy that level left, and the sums run in
// the same order as above (ascending columns), so the results are the same
// bits.  LDS holds two levels of kTriLds values.
constexpr int kTriLds = 2048;
constexpr int kTriNone = std::numeric_limits<int>::max();
//
// ALL_LDS (every dependency is on the previous level, so the sweep never
// reads its own global output): the barrier between levels orders LDS only —
// the output stores stay in flight — and each lane's first row of the levels
// kTriAhead ahead is loaded into a register ring (structure, row id, the
// input taken in level order from `in_pos`, 1/U(i,i)) while the current level
// computes, so a level costs LDS latency and a barrier rather than an HBM
// round trip.
constexpr int kTriAhead = 3;
template <int K, class V, class T, bool UPPER>
__global__ void __launch_bounds__(kTriThreads)
k_tri_ell(int n_levels, const int* __restrict__ lptr, const int* __restrict__ order, const int* __restrict__ dep,
          const V* __restrict__ val, const T* __restrict__ invd, const T* __restrict__ in, T* out, int use_lds,
          const int* skip) {
    __shared__ T buf[2 * kTriLds];
    if (skip && *skip) return;
    for (int lev = 0; lev < n_levels; ++lev) {
        const int a = lptr[lev], b = lptr[lev + 1];
        T* cur = buf + (lev & 1) * kTriLds;
        const T* prev = buf + ((lev & 1) ^ 1) * kTriLds;
        for (int idx = a + (int)threadIdx.x; idx < b; idx += kTriThreads) {
            int d[K];
            V v[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                d[k] = dep[(size_t)idx * K + k];
                v[k] = val[(size_t)idx * K + k];
            }
            const int i = order[idx];
            T s = in[i];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (d[k] == kTriNone) continue;
                const T yj = d[k] < 0 ? prev[-d[k] - 1] : out[d[k]];
                s -= (T)v[k] * yj;
            }
            if (UPPER) s = s * invd[idx];
            out[i] = s;
            if (use_lds) cur[idx - a] = s;
        }
        __syncthreads();
    }
}

template <int K, class V, class T>
struct TriRow {
    int d[K];
    V v[K];
    int i;
    T x, dv;
};

template <int K, class V, class T, bool UPPER>
__global__ void __launch_bounds__(kTriThreads)
k_tri_ell_lds(int n_levels, const int* __restrict__ lptr, const int* __restrict__ order, const int* __restrict__ dep,
              const V* __restrict__ val, const T* __restrict__ invd, const T* __restrict__ in_pos, T* out,
              const int* skip) {
    __shared__ T buf[2 * kTriLds];
    if (skip && *skip) return;
    const int tid = threadIdx.x;
    TriRow<K, V, T> ring[kTriAhead];
    auto load = [&](TriRow<K, V, T>& R, int idx) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            R.d[k] = dep[(size_t)idx * K + k];
            R.v[k] = val[(size_t)idx * K + k];
        }
        R.i = order[idx];
        R.x = in_pos[idx];
        if (UPPER) R.dv = invd[idx];
    };
    auto solve_row = [&](const TriRow<K, V, T>& R, const T* prev, T* cur, int slot) {
        T s = R.x;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (R.d[k] != kTriNone) s -= (T)R.v[k] * prev[-R.d[k] - 1];
        if (UPPER) s = s * R.dv;
        out[R.i] = s;
        cur[slot] = s;
    };
#pragma unroll
    for (int r = 0; r < kTriAhead; ++r)
        if (r < n_levels && lptr[r] + tid < lptr[r + 1]) load(ring[r], lptr[r] + tid);
    for (int lev0 = 0; lev0 < n_levels; lev0 += kTriAhead) {
#pragma unroll
        for (int r = 0; r < kTriAhead; ++r) {  // constant ring slots
            const int lev = lev0 + r;
            if (lev >= n_levels) break;
            const int a = lptr[lev], b = lptr[lev + 1];
            T* cur = buf + (lev & 1) * kTriLds;
            const T* prev = buf + ((lev & 1) ^ 1) * kTriLds;
            if (a + tid < b) solve_row(ring[r], prev, cur, tid);
            for (int idx = a + tid + kTriThreads; idx < b; idx += kTriThreads) {  // levels wider than the workgroup
                TriRow<K, V, T> R;
                load(R, idx);
                solve_row(R, prev, cur, idx - a);
            }
            const int nl = lev + kTriAhead;
            if (nl < n_levels && lptr[nl] + tid < lptr[nl + 1]) load(ring[r], lptr[nl] + tid);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }
    }
}

// ALL_LDS sweep whose levels are all at most ROWS * kTriThreads rows wide (a
// 2-D grid up to 1024 / 2048 wide in natural order; two rows per lane only for
// factor rows of at most two entries): lane t owns rows t and t + 1024 of
// every level, and the
// level loop holds no branch that depends on the lane or on a loop of its own,
// so the ring loads several levels ahead stay in flight across the barriers
// instead of being waited for at the top of each level.  The result goes out
// in level order (out_pos[position], one contiguous run per level instead of
// one cache line per row); k_tri_scatter puts it in row order afterwards.  A
// lane past the end of its level loads and solves the level's last row again
// (the same bits as the owner lane: its duplicate store writes the value
// already there) and parks its LDS copy in a slot the next level does not
// read.  A missing dependency (kTriNone) reads slot 0 and discards it with a
// select; the subtraction order is the one of k_tri_ell_lds.
template <int K, class V, class T, bool UPPER, int ROWS>
__global__ void __launch_bounds__(kTriThreads)
k_tri_ell_narrow(int n_levels, const int* __restrict__ lptr, const int* __restrict__ dep,
                 const V* __restrict__ val, const T* __restrict__ invd, const T* __restrict__ in_pos,
                 T* __restrict__ out_pos, const int* skip) {
    static_assert(ROWS == 1 || (ROWS == 2 && K <= 2), "two rows per lane only for two-entry factor rows");
    static_assert(ROWS * kTriThreads <= kTriLds, "a level's LDS copy holds every position");
    // levels each lane loads ahead: for factors with at most 2 entries per
    // row (five-point ILU(0)) measured on MI355X (conv-diff 1024^2,
    // FGMRES(30) + ILU(0)): FP32 3 -> 12 levels, mixed solve 43.6 -> 37.9 ms;
    // FP64 best at 8 (67.3 -> 63.4 ms; 12 levels: 91 ms); half as many with
    // two rows per lane (the same rows in flight).  Wider rows keep 3 (a
    // deeper ring would not fit the 128 registers of a 1024-thread workgroup).
    constexpr int kAhead = K <= 2 ? (sizeof(T) == 4 ? 12 : 8) / ROWS : kTriAhead;
    __shared__ T buf[2 * kTriLds];
    if ((skip && *skip) || n_levels <= 0) return;
    const int tid = threadIdx.x;
    struct Rows {
        TriRow<K, V, T> q[ROWS];
    };
    Rows ring[kAhead];
    // lane tid owns positions tid + q * kTriThreads of every level; levels are
    // never empty, so b - 1 is a row of level lev
    auto load = [&](Rows& R, int lev) {
        const int a = lptr[lev], b = lptr[lev + 1];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int idx = min(a + tid + q * kTriThreads, b - 1);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                R.q[q].d[k] = dep[(size_t)idx * K + k];
                R.q[q].v[k] = val[(size_t)idx * K + k];
            }
            R.q[q].i = idx;  // position in level order
            R.q[q].x = in_pos[idx];
            if (UPPER) R.q[q].dv = invd[idx];
        }
    };
    const int last = n_levels - 1;
    auto step = [&](Rows& R, int lev) {
        T* cur = buf + (lev & 1) * kTriLds;
        const T* prev = buf + ((lev & 1) ^ 1) * kTriLds;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            T s = R.q[q].x;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const bool has = R.q[q].d[k] != kTriNone;
                const T yj = prev[has ? -R.q[q].d[k] - 1 : 0];
                const T t = s - (T)R.q[q].v[k] * yj;
                s = has ? t : s;
            }
            if (UPPER) s = s * R.q[q].dv;
            out_pos[R.q[q].i] = s;
            cur[tid + q * kTriThreads] = s;
        }
        load(R, min(lev + kAhead, last));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };
#pragma unroll
    for (int r = 0; r < kAhead; ++r) load(ring[r], min(r, last));
    // whole rounds of kAhead levels with no exit inside (every path into
    // the loop header then has the same loads in flight), the rest after
    const int whole = n_levels - n_levels % kAhead;
    for (int lev0 = 0; lev0 < whole; lev0 += kAhead) {
#pragma unroll
        for (int r = 0; r < kAhead; ++r) step(ring[r], lev0 + r);  // constant ring slots
    }
#pragma unroll
    for (int r = 0; r < kAhead - 1; ++r)
        if (whole + r < n_levels) step(ring[r], whole + r);
}

// out[order[p]] = out_pos[p]: a narrow sweep's result back in row order.
template <class T>
__global__ void k_tri_scatter(int n, const int* __restrict__ order, const T* __restrict__ out_pos, T* __restrict__ out,
                              const int* skip) {
    if (skip && *skip) return;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) out[order[p]] = out_pos[p];
}

// in_pos[p] = in[order[p]]: a sweep's input in level order.
template <class T>
__global__ void k_tri_gather(int n, const int* __restrict__ order, const T* __restrict__ in, T* __restrict__ in_pos,
                             const int* skip) {
    if (skip && *skip) return;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) in_pos[p] = in[order[p]];
}

// ---- helpers ---------------------------------------------------------------------
static DevBuf upload(Ctx* c, const void* h, size_t bytes) {
    DevBuf b(c, bytes);
    upload_on(c->stream, b.p, h, bytes);
    return b;
}
static DevBuf upload_f32(Ctx* c, const std::vector<double>& v) {
    std::vector<float> f(v.begin(), v.end());
    return upload(c, f.data(), f.size() * sizeof(float));
}
static hsps_host::HostCsr to_host(const Csr& A) {
    hsps_host::HostCsr h;
    h.rows = A.rows;
    h.cols = A.cols;
    h.rp = A.h_rp;
    h.ci = A.h_ci;
    h.v = A.h_v;
    return h;
}
static std::unique_ptr<Csr> to_dev(Ctx* c, hsps_host::HostCsr&& h) {
    return make_csr(c, h.rows, h.cols, std::move(h.rp), std::move(h.ci), std::move(h.v), true);
}

static TriSchedule make_schedule(Ctx* c, const std::vector<int>& lev, int n_levels) {
    TriSchedule s;
    const int n = (int)lev.size();
    s.n_levels = n_levels;
    s.h_level_ptr.assign(n_levels + 1, 0);
    for (int i = 0; i < n; ++i) s.h_level_ptr[lev[i] + 1]++;
    for (int l = 0; l < n_levels; ++l) {
        s.max_level_width = std::max(s.max_level_width, s.h_level_ptr[l + 1]);
        s.h_level_ptr[l + 1] += s.h_level_ptr[l];
    }
    std::vector<int> order(n), pos(s.h_level_ptr.begin(), s.h_level_ptr.end() - 1);
    for (int i = 0; i < n; ++i) order[pos[lev[i]]++] = i;
    s.h_order = order;
    s.order = upload(c, order.data(), order.size() * sizeof(int));
    s.level_ptr = upload(c, s.h_level_ptr.data(), s.h_level_ptr.size() * sizeof(int));
    return s;
}

// ---- setup -------------------------------------------------------------------------
static void setup_jacobi(Precond& M, const Csr& A, double omega) {
    std::vector<double> d(A.rows);
    for (int i = 0; i < A.rows; ++i) {
        double a = 0.0;
        for (int k = A.h_rp[i]; k < A.h_rp[i + 1]; ++k)
            if (A.h_ci[k] == i) { a = A.h_v[k]; break; }
        if (std::abs(a) < kEps)
            throw Error(HSPS_ERR_ZERO_DIAG, "JacobiPreconditioner: zero diagonal at row " + std::to_string(i));
        d[i] = omega / a;
    }
    M.dinv64 = upload(M.ctx, d.data(), d.size() * sizeof(double));
    M.dinv32 = upload_f32(M.ctx, d);
    M.jacobi_omega = omega;
}

// Level-ordered ELL of the entries a sweep uses (lower: j < i; upper: j > i),
// ascending columns per row as the CSR scan visits them.
static TriEll make_tri_ell(Ctx* c, const hsps_host::HostCsr& F, const std::vector<int>& lev, const TriSchedule& S,
                           bool lower, const std::vector<double>& invd) {
    TriEll E;
    const int n = F.rows;
    int maxc = 0;
    for (int i = 0; i < n; ++i) {
        int cnt = 0;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) cnt += lower ? F.ci[k] < i : F.ci[k] > i;
        maxc = std::max(maxc, cnt);
    }
    if (maxc > 8) return E;  // the CSR sweep serves such factors
    const int K = maxc <= 2 ? 2 : maxc <= 4 ? 4 : 8;
    E.K = K;
    E.lds = S.max_level_width <= kTriLds;
    std::vector<int> pos_of(n);
    for (int p = 0; p < n; ++p) pos_of[S.h_order[p]] = p;
    std::vector<int> dep((size_t)n * K, kTriNone);
    std::vector<double> val((size_t)n * K, 0.0), ip(lower ? 0 : n);
    bool all_lds = E.lds;
    for (int p = 0; p < n; ++p) {
        const int i = S.h_order[p], li = lev[i];
        int c = 0;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {
            const int j = F.ci[k];
            if (lower ? j >= i : j <= i) continue;
            const bool from_lds = E.lds && li > 0 && lev[j] == li - 1;
            dep[(size_t)p * K + c] = from_lds ? -(1 + pos_of[j] - S.h_level_ptr[li - 1]) : j;
            all_lds = all_lds && from_lds;
            val[(size_t)p * K + c] = F.v[k];
            ++c;
        }
        if (!lower) ip[p] = invd[i];
    }
    E.all_lds = all_lds;
    E.dep = upload(c, dep.data(), dep.size() * sizeof(int));
    E.val64 = upload(c, val.data(), val.size() * sizeof(double));
    E.val32 = upload_f32(c, val);
    if (!lower) {
        E.invd64 = upload(c, ip.data(), ip.size() * sizeof(double));
        E.invd32 = upload_f32(c, ip);
    }
    return E;
}

static void setup_ilu(Precond& M, const Csr& A, double drop_tol, int fill_p) {
    hsps_host::HostCsr F;
    try {
        F = (drop_tol > 0.0 || fill_p > 0) ? hsps_host::ilut(to_host(A), drop_tol, fill_p) : hsps_host::ilu0(to_host(A));
    } catch (const std::runtime_error& e) {
        throw Error(HSPS_ERR_PIVOT, e.what());
    }
    auto f = std::make_unique<IluFactor>();
    // U(i,i) as the backward sweep reads it: 1.0 when row i has no diagonal
    // entry; a near-zero value is an error (ilu.cpp:250-259), raised here at
    // setup instead of at the first apply.
    std::vector<double> invd(F.rows, 1.0);
    for (int i = 0; i < F.rows; ++i) {
        double d = 1.0;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k)
            if (F.ci[k] == i) { d = F.v[k]; break; }
        if (std::abs(d) < kEps)
            throw Error(HSPS_ERR_PIVOT, "ILU backward sub: near-zero U diagonal at row " + std::to_string(i));
        invd[i] = 1.0 / d;
    }
    f->tmp64 = DevBuf(M.ctx, (size_t)F.rows * sizeof(double));
    f->tmp32 = DevBuf(M.ctx, (size_t)F.rows * sizeof(float));
    f->fwd_band = make_tri_band(M.ctx, F, true, invd);
    f->bwd_band = make_tri_band(M.ctx, F, false, invd);
    f->band_err = DevBuf(M.ctx, sizeof(int));
    fill_on(M.ctx->stream, f->band_err.p, 0, sizeof(int));
    f->l_empty = 1;
    for (int i = 0; i < F.rows && f->l_empty; ++i)
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k)
            if (F.ci[k] < i) {
                f->l_empty = 0;
                break;
            }
    // the level-scheduled sweeps (and the factor on the device) only when a
    // sweep with work is not taken by a band layout, which holds everything
    // its sweep reads
    if (!((f->fwd_band.ok || f->l_empty) && f->bwd_band.ok)) {
        int nl_lo = 0, nl_up = 0;
        auto lo = hsps_host::tri_levels(F, true, &nl_lo);
        auto up = hsps_host::tri_levels(F, false, &nl_up);
        f->fwd = make_schedule(M.ctx, lo, nl_lo);
        f->bwd = make_schedule(M.ctx, up, nl_up);
        f->fwd_ell = make_tri_ell(M.ctx, F, lo, f->fwd, true, invd);
        f->bwd_ell = make_tri_ell(M.ctx, F, up, f->bwd, false, invd);
        f->inv_diag64 = upload(M.ctx, invd.data(), invd.size() * sizeof(double));
        f->inv_diag32 = upload_f32(M.ctx, invd);
        f->pos64 = DevBuf(M.ctx, (size_t)F.rows * sizeof(double));
        f->pos32 = DevBuf(M.ctx, (size_t)F.rows * sizeof(float));
        f->opos64 = DevBuf(M.ctx, (size_t)F.rows * sizeof(double));
        f->opos32 = DevBuf(M.ctx, (size_t)F.rows * sizeof(float));
        f->lu = to_dev(M.ctx, std::move(F));
    }
    M.ilu = std::move(f);
}

void precond_check(const Precond* M) {
    if (!M || M->kind != HSPS_PRECOND_ILU || !M->ilu || !(M->ilu->fwd_band.ok || M->ilu->bwd_band.ok)) return;
    int e = 0;
    HSPS_HIP(hipMemcpyAsync(&e, M->ilu->band_err.p, sizeof(int), hipMemcpyDeviceToHost, M->ctx->stream));
    HSPS_HIP(hipStreamSynchronize(M->ctx->stream));
    if (e) {
        fill_on(M->ctx->stream, M->ilu->band_err.p, 0, sizeof(int));
        throw Error(HSPS_ERR_HIP, "ILU band sweep: a wait between waves gave up");
    }
}

template <int K, class T, bool UPPER>
static void tri_ell_launch(const TriSchedule& S, const TriEll& E, const T* in, T* out, T* in_pos, T* out_pos,
                           hipStream_t s, const int* skip) {
    const bool dbl = std::is_same<T, double>::value;
    const T* invd = UPPER ? (dbl ? (const T*)E.invd64.p : (const T*)E.invd32.p) : nullptr;
    const int n = S.n_levels ? S.h_level_ptr.back() : 0;
    if (E.all_lds) {
        hipLaunchKernelGGL(k_tri_gather<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, S.order.as<int>(), in, in_pos,
                           skip);
        const bool one = S.max_level_width <= kTriThreads;
        if (one || (K <= 2 && S.max_level_width <= 2 * kTriThreads)) {
            auto narrow = [&](auto rows) {
                constexpr int R = decltype(rows)::value;
                if constexpr (std::is_same<T, double>::value)
                    hipLaunchKernelGGL((k_tri_ell_narrow<K, double, T, UPPER, R>), dim3(1), dim3(kTriThreads), 0, s,
                                       S.n_levels, S.level_ptr.as<int>(), E.dep.as<int>(), E.val64.as<double>(), invd,
                                       (const T*)in_pos, out_pos, skip);
                else
                    hipLaunchKernelGGL((k_tri_ell_narrow<K, float, T, UPPER, R>), dim3(1), dim3(kTriThreads), 0, s,
                                       S.n_levels, S.level_ptr.as<int>(), E.dep.as<int>(), E.val32.as<float>(), invd,
                                       (const T*)in_pos, out_pos, skip);
            };
            if (one) narrow(std::integral_constant<int, 1>{});
            else if constexpr (K <= 2) narrow(std::integral_constant<int, 2>{});
            HSPS_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_tri_scatter<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, S.order.as<int>(),
                               (const T*)out_pos, out, skip);
        } else if (dbl)
            hipLaunchKernelGGL((k_tri_ell_lds<K, double, T, UPPER>), dim3(1), dim3(kTriThreads), 0, s, S.n_levels,
                               S.level_ptr.as<int>(), S.order.as<int>(), E.dep.as<int>(), E.val64.as<double>(), invd,
                               (const T*)in_pos, out, skip);
        else
            hipLaunchKernelGGL((k_tri_ell_lds<K, float, T, UPPER>), dim3(1), dim3(kTriThreads), 0, s, S.n_levels,
                               S.level_ptr.as<int>(), S.order.as<int>(), E.dep.as<int>(), E.val32.as<float>(), invd,
                               (const T*)in_pos, out, skip);
    } else if (dbl) {
        hipLaunchKernelGGL((k_tri_ell<K, double, T, UPPER>), dim3(1), dim3(kTriThreads), 0, s, S.n_levels,
                           S.level_ptr.as<int>(), S.order.as<int>(), E.dep.as<int>(), E.val64.as<double>(), invd, in,
                           out, E.lds, skip);
    } else {
        hipLaunchKernelGGL((k_tri_ell<K, float, T, UPPER>), dim3(1), dim3(kTriThreads), 0, s, S.n_levels,
                           S.level_ptr.as<int>(), S.order.as<int>(), E.dep.as<int>(), E.val32.as<float>(), invd, in,
                           out, E.lds, skip);
    }
    HSPS_CHECK_LAUNCH();
}
template <class T, bool UPPER>
static void tri_ell_sweep(const TriSchedule& S, const TriEll& E, const T* in, T* out, T* in_pos, T* out_pos,
                          hipStream_t s, const int* skip) {
    if (E.K == 2) tri_ell_launch<2, T, UPPER>(S, E, in, out, in_pos, out_pos, s, skip);
    else if (E.K == 4) tri_ell_launch<4, T, UPPER>(S, E, in, out, in_pos, out_pos, s, skip);
    else tri_ell_launch<8, T, UPPER>(S, E, in, out, in_pos, out_pos, s, skip);
}

static void alloc_level_work(Ctx* c, AmgLevelDev& L, int n) {
    L.x64 = DevBuf(c, (size_t)n * 8);
    L.x64b = DevBuf(c, (size_t)n * 8);
    L.rhs64 = DevBuf(c, (size_t)n * 8);
    L.res64 = DevBuf(c, (size_t)n * 8);
    L.x32 = DevBuf(c, (size_t)n * 4);
    L.x32b = DevBuf(c, (size_t)n * 4);
    L.rhs32 = DevBuf(c, (size_t)n * 4);
    L.res32 = DevBuf(c, (size_t)n * 4);
}

// Host copy of a CSR: the arrays kept at creation, or a download for levels
// built on the device.
static hsps_host::HostCsr host_copy(const Csr& A) {
    if (A.h_ci.size() == (size_t)A.nnz && A.h_v.size() == (size_t)A.nnz) return to_host(A);
    hsps_host::HostCsr h;
    h.rows = A.rows;
    h.cols = A.cols;
    h.rp = A.h_rp;
    h.ci.resize(A.nnz);
    h.v.resize(A.nnz);
    A.ctx->set_current();
    if (A.nnz) {
        HSPS_HIP(hipMemcpyAsync(h.ci.data(), A.ci.p, (size_t)A.nnz * sizeof(int), hipMemcpyDeviceToHost, A.ctx->stream));
        HSPS_HIP(hipMemcpyAsync(h.v.data(), A.v64.p, (size_t)A.nnz * sizeof(double), hipMemcpyDeviceToHost,
                                A.ctx->stream));
    }
    HSPS_HIP(hipStreamSynchronize(A.ctx->stream));
    return h;
}

// Hierarchy on the device (amg_setup.hip), then the coarsest level's dense
// solve operator (amg.cpp:260-291), factored on the device (dense_lu.hip).
static void setup_amg(Precond& M, const Csr& A, const hsps_params& p) {
    M.amg_omega = p.amg_omega;
    M.pre = p.amg_smooth_pre;
    M.post = p.amg_smooth_post;
    M.amg_times = AmgSetupTimes{};
    amg_setup_device(M, A, p, &M.amg_times);
    const int L = (int)M.amg.size();
    for (int lv = 0; lv < L; ++lv) alloc_level_work(M.ctx, M.amg[lv], lv == 0 ? A.rows : M.amg[lv].A->rows);
    const auto t0 = std::chrono::steady_clock::now();
    const Csr& Ac = L == 1 ? A : *M.amg[L - 1].A;
    M.coarse_n = Ac.rows;
    dense_inverse_device(M.ctx, Ac, M.coarse_inv, &M.coarse_ld, M.ctx->stream);
    M.amg_times.coarse = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- apply -------------------------------------------------------------------------
template <class T> struct Lv;
template <> struct Lv<double> {
    static double* x(AmgLevelDev& L) { return L.x64.as<double>(); }
    static double* xb(AmgLevelDev& L) { return L.x64b.as<double>(); }
    static double* rhs(AmgLevelDev& L) { return L.rhs64.as<double>(); }
    static double* res(AmgLevelDev& L) { return L.res64.as<double>(); }
    static const double* dinv(const AmgLevelDev& L) { return L.dinv64.as<double>(); }
};
template <> struct Lv<float> {
    static float* x(AmgLevelDev& L) { return L.x32.as<float>(); }
    static float* xb(AmgLevelDev& L) { return L.x32b.as<float>(); }
    static float* rhs(AmgLevelDev& L) { return L.rhs32.as<float>(); }
    static float* res(AmgLevelDev& L) { return L.res32.as<float>(); }
    static const float* dinv(const AmgLevelDev& L) { return L.dinv32.as<float>(); }
};

static void smooth(const Csr& A, const double* d, double w, const double* rhs, const double* xi, double* xo,
                   hipStream_t s, const int* skip) {
    smooth_f64(A, d, w, rhs, xi, xo, s, skip);
}
static void smooth(const Csr& A, const float* d, double w, const float* rhs, const float* xi, float* xo,
                   hipStream_t s, const int* skip) {
    smooth_f32(A, d, (float)w, rhs, xi, xo, s, skip);
}
static void resid(const Csr& A, const double* b, const double* x, double* r, hipStream_t s, const int* skip) {
    residual_f64(A, b, x, r, s, nullptr, nullptr, skip);
}
static void resid(const Csr& A, const float* b, const float* x, float* r, hipStream_t s, const int* skip) {
    residual_f32(A, b, x, r, s, skip);
}
static void fill_zero(int n, double* x, hipStream_t s, const int* skip) { fill_zero_f64(n, x, s, skip); }
static void fill_zero(int n, float* x, hipStream_t s, const int* skip) { fill_zero_f32(n, x, s, skip); }
static void mv(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip) { spmv_f64(A, x, y, s, skip); }
static void mv(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip) { spmv_f32(A, x, y, s, skip); }
static void mv_add(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip) {
    spmv_add_f64(A, x, y, s, skip);
}
static void mv_add(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip) {
    spmv_add_f32(A, x, y, s, skip);
}

// V-cycle (amg.cpp:296-335) on level lv: rhs -> x (x starts at zero).  `out`
// is where the level's result must land (level 0: the caller's z).
template <class T>
static void vcycle(const Precond& Mc, const Csr& A0, int lv, const T* rhs, T* out, hipStream_t s, const int* skip) {
    Precond& M = const_cast<Precond&>(Mc);
    const int L = (int)M.amg.size();
    AmgLevelDev& Lv_ = M.amg[lv];
    const Csr& A = lv == 0 ? A0 : *Lv_.A;
    const int n = A.rows;
    if (lv == L - 1) {
        hipLaunchKernelGGL(k_dense_gemv<T>, dim3((n + 3) / 4), dim3(kBlock), 0, s, n, M.coarse_ld,
                           M.coarse_inv.as<double>(), rhs, out, skip);
        HSPS_CHECK_LAUNCH();
        return;
    }
    const T* dinv = Lv<T>::dinv(Lv_);
    // out-of-place sweeps after the first: (pre-1) + post; land the last in out
    const int sweeps = (M.pre > 0 ? M.pre - 1 : 0) + M.post;
    T* cur = (sweeps % 2 == 0) ? out : Lv<T>::xb(Lv_);
    T* oth = (cur == out) ? Lv<T>::xb(Lv_) : out;
    if (M.pre > 0) {
        hipLaunchKernelGGL(k_first_sweep<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, (T)M.amg_omega, dinv, rhs, cur,
                           skip);
        HSPS_CHECK_LAUNCH();
        for (int k = 1; k < M.pre; ++k) {
            smooth(A, dinv, M.amg_omega, rhs, cur, oth, s, skip);
            std::swap(cur, oth);
        }
    } else {
        fill_zero(n, cur, s, skip);
    }
    T* res = Lv<T>::res(Lv_);
    resid(A, rhs, cur, res, s, skip);
    AmgLevelDev& Cn = M.amg[lv + 1];
    T* rc = Lv<T>::rhs(Cn);
    T* ec = Lv<T>::x(Cn);
    mv(*Lv_.R, res, rc, s, skip);
    vcycle<T>(M, A0, lv + 1, rc, ec, s, skip);
    mv_add(*Lv_.P, ec, cur, s, skip);
    for (int k = 0; k < M.post; ++k) {
        smooth(A, dinv, M.amg_omega, rhs, cur, oth, s, skip);
        std::swap(cur, oth);
    }
}

template <class T>
static void apply_t(const Precond& M, const T* r, T* z, hipStream_t s, const int* skip) {
    const int n = M.n;
    switch (M.kind) {
        case HSPS_PRECOND_NONE:
            hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, z, skip);
            HSPS_CHECK_LAUNCH();
            break;
        case HSPS_PRECOND_JACOBI: {
            const T* d = std::is_same<T, double>::value ? (const T*)M.dinv64.p : (const T*)M.dinv32.p;
            hipLaunchKernelGGL(k_diag_scale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, r, z, skip);
            HSPS_CHECK_LAUNCH();
            break;
        }
        case HSPS_PRECOND_ILU: {
            const IluFactor& f = *M.ilu;
            const bool dbl = std::is_same<T, double>::value;
            T* y = dbl ? (T*)f.tmp64.p : (T*)f.tmp32.p;
            const T* invd = dbl ? (const T*)f.inv_diag64.p : (const T*)f.inv_diag32.p;
            if ((f.fwd_band.ok || f.l_empty) && f.bwd_band.ok) {
                int* err = f.band_err.as<int>();
                const T* yin = r;  // unit L with nothing below the diagonal: y = r exactly
                if (!f.l_empty) {
                    tri_band_sweep<T>(f.fwd_band, false, n, r, y, err, s, skip);
                    yin = y;
                }
                tri_band_sweep<T>(f.bwd_band, true, n, yin, z, err, s, skip);
            } else if (f.fwd_ell.K && f.bwd_ell.K) {
                T* ip = dbl ? (T*)f.pos64.p : (T*)f.pos32.p;
                T* op = dbl ? (T*)f.opos64.p : (T*)f.opos32.p;
                int* err = f.band_err.as<int>();
                if (f.fwd_band.ok) tri_band_sweep<T>(f.fwd_band, false, n, r, y, err, s, skip);
                else tri_ell_sweep<T, false>(f.fwd, f.fwd_ell, r, y, ip, op, s, skip);
                if (f.bwd_band.ok) tri_band_sweep<T>(f.bwd_band, true, n, (const T*)y, z, err, s, skip);
                else tri_ell_sweep<T, true>(f.bwd, f.bwd_ell, (const T*)y, z, ip, op, s, skip);
            } else if (dbl) {
                hipLaunchKernelGGL((k_tri_lower<double, T>), dim3(1), dim3(kTriThreads), 0, s, f.fwd.n_levels,
                                   f.fwd.level_ptr.as<int>(), f.fwd.order.as<int>(), f.lu->rp.as<int>(),
                                   f.lu->ci.as<int>(), f.lu->v64.as<double>(), r, y, skip);
                HSPS_CHECK_LAUNCH();
                hipLaunchKernelGGL((k_tri_upper<double, T>), dim3(1), dim3(kTriThreads), 0, s, f.bwd.n_levels,
                                   f.bwd.level_ptr.as<int>(), f.bwd.order.as<int>(), f.lu->rp.as<int>(),
                                   f.lu->ci.as<int>(), f.lu->v64.as<double>(), invd, (const T*)y, z, skip);
                HSPS_CHECK_LAUNCH();
            } else {
                hipLaunchKernelGGL((k_tri_lower<float, T>), dim3(1), dim3(kTriThreads), 0, s, f.fwd.n_levels,
                                   f.fwd.level_ptr.as<int>(), f.fwd.order.as<int>(), f.lu->rp.as<int>(),
                                   f.lu->ci.as<int>(), f.lu->v32.as<float>(), r, y, skip);
                HSPS_CHECK_LAUNCH();
                hipLaunchKernelGGL((k_tri_upper<float, T>), dim3(1), dim3(kTriThreads), 0, s, f.bwd.n_levels,
                                   f.bwd.level_ptr.as<int>(), f.bwd.order.as<int>(), f.lu->rp.as<int>(),
                                   f.lu->ci.as<int>(), f.lu->v32.as<float>(), invd, (const T*)y, z, skip);
                HSPS_CHECK_LAUNCH();
            }
            break;
        }
        case HSPS_PRECOND_AMG:
            // level 0 smooths with the caller's matrix; the coarse levels and
            // D^-1 were built from the values it had at setup
            if (M.A0->values_version != M.A0_version)
                throw Error(HSPS_ERR_INVALID, "AMG preconditioner: the matrix values changed since setup "
                                              "(hsps_csr_update_values); create the preconditioner again");
            vcycle<T>(M, *M.A0, 0, r, z, s, skip);
            break;
        case HSPS_PRECOND_USER: {  // the caller's apply (preconditioner_base.hpp:32), FP64 vectors
            int rc;
            if constexpr (std::is_same<T, double>::value) {
                rc = M.user_fn(M.user_data, r, z, (void*)s);
            } else {
                double* r64 = M.user_r64.as<double>();
                double* z64 = M.user_z64.as<double>();
                hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, r64, skip);
                HSPS_CHECK_LAUNCH();
                rc = M.user_fn(M.user_data, r64, z64, (void*)s);
                hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(kBlock), 0, s, n,
                                   (const double*)z64, z, skip);
                HSPS_CHECK_LAUNCH();
            }
            if (rc < 0) throw Error(rc, "user preconditioner: apply returned status " + std::to_string(rc));
            break;
        }
        default:
            throw Error(HSPS_ERR_INVALID, "unknown preconditioner kind");
    }
}

void precond_apply_f64(const Precond& M, const double* r, double* z, hipStream_t s, const int* skip) {
    apply_t<double>(M, r, z, s, skip);
}
void precond_apply_f32(const Precond& M, const float* r, float* z, hipStream_t s, const int* skip) {
    apply_t<float>(M, r, z, s, skip);
}
int precond_kind(const Precond& M) { return M.kind; }
const double* precond_jacobi_dinv64(const Precond& M) { return M.dinv64.as<double>(); }
const float* precond_jacobi_dinv32(const Precond& M) { return M.dinv32.as<float>(); }

}  // namespace hsps_dev

using namespace hsps_dev;

extern "C" {

int hsps_precond_create(hsps_csr A, int kind, const hsps_params* pp, hsps_precond* out, double* setup_seconds) {
    return guarded([&] {
        if (!A || !out) throw Error(HSPS_ERR_INVALID, "hsps_precond_create: null argument");
        if (A->rows != A->cols) throw Error(HSPS_ERR_SIZE, "preconditioner needs a square matrix");
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        const auto t0 = std::chrono::steady_clock::now();
        A->ctx->set_current();
        auto* M = new hsps_precond_s;
        M->kind = kind;
        M->ctx = A->ctx;
        M->n = A->rows;
        M->A0 = A;
        M->A0_version = A->values_version;
        try {
            switch (kind) {
                case HSPS_PRECOND_NONE: break;
                case HSPS_PRECOND_JACOBI: setup_jacobi(*M, *A, p.jacobi_omega); break;
                case HSPS_PRECOND_ILU: setup_ilu(*M, *A, p.drop_tol, p.fill_per_row); break;
                case HSPS_PRECOND_AMG: setup_amg(*M, *A, p); break;
                default: throw Error(HSPS_ERR_INVALID, "SolverFactory: unknown PrecondType");
            }
            HSPS_HIP(hipStreamSynchronize(A->ctx->stream));
        } catch (...) {
            delete M;
            throw;
        }
        if (setup_seconds)
            *setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ctx_retain(M->ctx);
        csr_retain(A);
        *out = M;
        return HSPS_OK;
    });
}

int hsps_ilu_factor(int n, const int* rp, const int* ci, const double* v, double drop_tol, int fill_per_row,
                    int* nnz_out, int* lu_rp, int* lu_ci, double* lu_v) {
    return guarded([&] {
        if (n < 0 || !rp || !nnz_out || (rp[n] > 0 && (!ci || !v)))
            throw Error(HSPS_ERR_INVALID, "hsps_ilu_factor: bad argument");
        hsps_host::HostCsr A;
        A.rows = A.cols = n;
        A.rp.assign(rp, rp + n + 1);
        A.ci.assign(ci, ci + rp[n]);
        A.v.assign(v, v + rp[n]);
        hsps_host::HostCsr F;
        try {
            F = (drop_tol > 0.0 || fill_per_row > 0) ? hsps_host::ilut(A, drop_tol, fill_per_row) : hsps_host::ilu0(A);
        } catch (const std::runtime_error& e) {
            throw Error(HSPS_ERR_PIVOT, e.what());
        }
        *nnz_out = F.nnz();
        if (lu_rp) std::copy(F.rp.begin(), F.rp.end(), lu_rp);
        if (lu_ci && lu_v) {
            std::copy(F.ci.begin(), F.ci.end(), lu_ci);
            std::copy(F.v.begin(), F.v.end(), lu_v);
        }
        return HSPS_OK;
    });
}

int hsps_precond_create_user(hsps_ctx ctx, int n, hsps_user_apply_fn apply, void* user, hsps_precond* out) {
    return guarded([&] {
        if (!ctx || !apply || !out || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_precond_create_user: bad argument");
        ctx->set_current();
        auto* M = new hsps_precond_s;
        M->kind = HSPS_PRECOND_USER;
        M->ctx = ctx;
        M->n = n;
        M->user_fn = apply;
        M->user_data = user;
        try {
            M->user_r64 = DevBuf(ctx, (size_t)std::max(n, 1) * sizeof(double));
            M->user_z64 = DevBuf(ctx, (size_t)std::max(n, 1) * sizeof(double));
        } catch (...) {
            delete M;
            throw;
        }
        ctx_retain(ctx);
        *out = M;
        return HSPS_OK;
    });
}

int hsps_precond_destroy(hsps_precond M) {
    return guarded([&] {
        if (M) {
            Ctx* c = M->ctx;
            hsps_csr_s* A = M->A0;
            (void)hipSetDevice(c->device);
            (void)hipStreamSynchronize(c->stream);
            delete M;
            if (A) csr_release(A);
            ctx_release(c);
        }
        return HSPS_OK;
    });
}

int hsps_precond_kind(hsps_precond M) { return M ? M->kind : HSPS_ERR_INVALID; }

int hsps_precond_apply_f64(hsps_precond M, const double* r, double* z) {
    return guarded([&] {
        if (!M || !r || !z) throw Error(HSPS_ERR_INVALID, "hsps_precond_apply_f64: null argument");
        M->ctx->set_current();
        precond_apply_f64(*M, r, z, M->ctx->stream, nullptr);
        precond_check(M);
        return HSPS_OK;
    });
}

int hsps_amg_levels(hsps_precond M) {
    if (!M || M->kind != HSPS_PRECOND_AMG) return HSPS_ERR_INVALID;
    return (int)M->amg.size();
}

static const Csr& amg_A(const Precond& M, int lv) { return lv == 0 ? *M.A0 : *M.amg[lv].A; }

int hsps_amg_level_info(hsps_precond M, int lv, int* rows, int* nnz) {
    return guarded([&] {
        if (!M || M->kind != HSPS_PRECOND_AMG || lv < 0 || lv >= (int)M->amg.size())
            throw Error(HSPS_ERR_INVALID, "hsps_amg_level_info: bad level");
        const Csr& A = amg_A(*M, lv);
        if (rows) *rows = A.rows;
        if (nnz) *nnz = A.nnz;
        return HSPS_OK;
    });
}

int hsps_amg_aggregates(hsps_precond M, int lv, int* out) {
    return guarded([&] {
        if (!M || M->kind != HSPS_PRECOND_AMG || lv < 0 || lv + 1 >= (int)M->amg.size() || !out)
            throw Error(HSPS_ERR_INVALID, "hsps_amg_aggregates: bad level");
        const auto& a = M->amg[lv].agg;
        std::copy(a.begin(), a.end(), out);
        return HSPS_OK;
    });
}

int hsps_ilu_sweep_kinds(hsps_precond M, int* fwd, int* bwd, int* w_fwd, int* w_bwd) {
    if (!M || M->kind != HSPS_PRECOND_ILU || !M->ilu) return HSPS_ERR_INVALID;
    const IluFactor& f = *M->ilu;
    if (fwd) *fwd = f.fwd_band.ok ? 1 : f.l_empty && f.bwd_band.ok ? 2 : 0;  // as the apply chooses
    if (bwd) *bwd = f.bwd_band.ok;
    if (w_fwd) *w_fwd = f.fwd_band.ok ? f.fwd_band.W : 0;
    if (w_bwd) *w_bwd = f.bwd_band.ok ? f.bwd_band.W : 0;
    return HSPS_OK;
}

int hsps_amg_setup_profile(hsps_precond M, double* out) {
    if (!M || M->kind != HSPS_PRECOND_AMG || !out) return HSPS_ERR_INVALID;
    const AmgSetupTimes& t = M->amg_times;
    const double v[6] = {t.strength, t.aggregate, t.prolongation, t.transpose, t.galerkin, t.coarse};
    std::copy(v, v + 6, out);
    return HSPS_OK;
}

int hsps_amg_level_matrix(hsps_precond M, int lv, int which, int* rp, int* ci, double* v) {
    return guarded([&] {
        if (!M || M->kind != HSPS_PRECOND_AMG || lv < 0 || lv >= (int)M->amg.size())
            throw Error(HSPS_ERR_INVALID, "hsps_amg_level_matrix: bad level");
        if (which != 0 && lv + 1 >= (int)M->amg.size())
            throw Error(HSPS_ERR_INVALID, "hsps_amg_level_matrix: coarsest level has no P/R");
        const Csr& A = which == 0 ? amg_A(*M, lv) : which == 1 ? *M->amg[lv].P : *M->amg[lv].R;
        if (!rp) throw Error(HSPS_ERR_INVALID, "hsps_amg_level_matrix: null row_ptr");
        if (!ci || !v) {  // row pointer only (sizes the other arrays)
            std::copy(A.h_rp.begin(), A.h_rp.end(), rp);
            return HSPS_OK;
        }
        const hsps_host::HostCsr h = host_copy(A);
        std::copy(h.rp.begin(), h.rp.end(), rp);
        std::copy(h.ci.begin(), h.ci.end(), ci);
        std::copy(h.v.begin(), h.v.end(), v);
        return HSPS_OK;
    });
}

}  // extern "C"
