// =============================================================================
// batch.hip — batched PCG + Jacobi for many independent systems (no reference
// counterpart; the reference solves one system per call).
//
// Layout: the systems' CSRs are concatenated (block-diagonal, global column
// indices).  Each system is cut into tiles of <= tile_rows rows (and <= one
// LDS stage of nonzeros); one workgroup owns one tile.  Systems are sorted by
// size and grouped into waves whose working set fits the 256 MB Infinity
// Cache, so a wave's iterations run out of MALL instead of HBM.
//
// Per iteration three launches over the tiles of the active wave, each with
// per-system reductions finished by the system's last-arriving tile:
//   K1  q = A p and p.q  -> alpha          (or, in RESID state, the FP64 true
//                                           residual r = b - A x and r.z)
//   K2  x += alpha p, r -= alpha q, r.r, r.z -> rel, rho, beta, decisions
//   K3  p = D^-1 r + beta p                  (or, in FLUSH state, x64 += x_lo)
// Mixed precision: FP32 matrix values and Krylov vectors, FP64 reductions and
// scalars, x accumulated in FP64 through an FP32 low-order part that is
// flushed at every reliable update.  A system whose |r| drops below
// delta*max|r| (or below tol) goes FLUSH -> RESID -> PUPDATE: its FP64 true
// residual replaces r (and confirms convergence) without extra launches.
// The wave's iteration chunk is captured once in a hipGraph and replayed; the
// host polls the wave's states one chunk behind.
// =============================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "device_common.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

enum : int { BS_NORMAL = 0, BS_FLUSH = 1, BS_RESID = 2, BS_PUPDATE = 3 };
enum : int { BR_NONE = 0, BR_RELIABLE = 1, BR_CHECK = 2, BR_RESTART = 3 };

struct BState {
    double rho, rho_old, alpha, beta, bnorm, rel, rmax, true_rel;
    int iter, halt, done, flag, reason, reliable, restarts, conv;
};

struct BArgs {
    const int* rp;
    const int* ci;
    const double* v64;
    const float* v32;
    const double* dinv64;
    const float* dinv32;
    const int* tile_r0;
    const int* tile_r1;
    const int* tile_sys;
    const int* sys_t0;
    const int* sys_t1;
    BState* st;
    unsigned* cnt;
    double* partials;
    int T;  // total tiles
    double tol, delta;
    int max_iter;
};

// ---- per-system reduction: last-arriving tile of the system sums partials ------------
template <int NV>
__device__ __forceinline__ bool sys_sum(double (&v)[NV], const BArgs& B, int tile, int s, double (&tot)[NV]) {
    __shared__ double red[NV * (kBlock / 64)];
    __shared__ int is_last;
    block_sum<NV>(v, red);
    const int t0 = B.sys_t0[s], t1 = B.sys_t1[s];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) B.partials[(size_t)k * B.T + tile] = v[k];
        const unsigned t = __hip_atomic_fetch_add(&B.cnt[s], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == (unsigned)(t1 - t0 - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            acc[k] += __hip_atomic_load(&B.partials[(size_t)k * B.T + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    block_sum<NV>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) tot[k] = acc[k];
        B.cnt[s] = 0u;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool sys_ticket(const BArgs& B, int s) {
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(&B.cnt[s], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == (unsigned)(B.sys_t1[s] - B.sys_t0[s] - 1));
    }
    __syncthreads();
    if (is_last && threadIdx.x == 0) {
        B.cnt[s] = 0u;
        return true;
    }
    return false;
}

// Row sums of a tile (rows r0..r1) of A*x, LDS-staged; calls f(row, sum) once
// per row from one lane.  Same staging rule as the single-system SpMV.
template <class Acc, class V, class X, class F>
__device__ __forceinline__ void tile_spmv(int r0, int r1, const int* __restrict__ rp, const int* __restrict__ ci,
                                          const V* __restrict__ val, const X* __restrict__ x, Acc* prod, F&& f) {
    const int tid = threadIdx.x;
    const int k0 = rp[r0], k1 = rp[r1];
    if (k1 - k0 > kLdsNnz) {  // one long row
        Acc s = 0;
        for (int k = k0 + tid; k < k1; k += kBlock) s += (Acc)val[k] * (Acc)x[ci[k]];
        double v1[1] = {(double)s};
        __shared__ double lred[kBlock / 64];
        block_sum<1>(v1, lred);
        if (tid == 0) f(r0, (Acc)v1[0]);
        return;
    }
    for (int k = k0 + tid; k < k1; k += kBlock) prod[k - k0] = (Acc)val[k] * (Acc)x[ci[k]];
    __syncthreads();
    const int nrows = r1 - r0;
    int tpr = 1;
    while (tpr < 64 && tpr * 2 * nrows <= kBlock) tpr *= 2;
    const int g = tid / tpr, l = tid - g * tpr, groups = kBlock / tpr;
    for (int row = r0 + g; row < r1; row += groups) {
        const int a = rp[row] - k0, b = rp[row + 1] - k0;
        Acc s = 0;
        for (int k = a + l; k < b; k += tpr) s += prod[k];
        for (int o = tpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (l == 0) f(row, s);
    }
}

template <class T> struct Prec;
template <> struct Prec<double> {
    static __device__ const double* val(const BArgs& B) { return B.v64; }
    static __device__ const double* dinv(const BArgs& B) { return B.dinv64; }
};
template <> struct Prec<float> {
    static __device__ const float* val(const BArgs& B) { return B.v32; }
    static __device__ const float* dinv(const BArgs& B) { return B.dinv32; }
};

// ---- init: r64 = b - A x, |b|, |r| per system ----------------------------------------
__global__ void __launch_bounds__(kBlock)
k_b_init1(BArgs B, int t0, const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r64) {
    __shared__ double prod[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[2] = {0.0, 0.0};
    tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x, prod, [&](int i, double sum) {
        const double ri = b[i] - sum;
        r64[i] = ri;
        red[0] += ri * ri;
    });
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) red[1] += b[i] * b[i];
    double tot[2];
    if (!sys_sum<2>(red, B, tile, s, tot)) return;
    BState S{};
    S.bnorm = sqrt(tot[1]);
    const double r0n = sqrt(tot[0]);
    S.rmax = r0n;
    if (S.bnorm < kEps) {  // cg_solver.cpp:60-66
        S.done = S.conv = S.halt = 1;
        S.rel = r0n;
    } else if (r0n / S.bnorm < B.tol) {
        S.done = S.conv = S.halt = 1;
        S.rel = r0n / S.bnorm;
    } else {
        S.rel = r0n / S.bnorm;
        S.halt = B.max_iter <= 0 ? 1 : 0;
    }
    S.true_rel = S.rel;
    B.st[s] = S;
}

// r = (T) r64, x_lo = 0, p = D^-1 r, rho = r.z
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_init2(BArgs B, int t0, const double* __restrict__ r64, T* __restrict__ r, T* __restrict__ p,
          float* __restrict__ xlo) {
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    if (B.st[s].halt) return;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    const T* dinv = Prec<T>::dinv(B);
    double red[1] = {0.0};
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
        const T ri = (T)r64[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        p[i] = zi;
        if (MIXED) xlo[i] = 0.0f;
        red[0] += (double)ri * (double)zi;
    }
    double tot[1];
    if (!sys_sum<1>(red, B, tile, s, tot)) return;
    B.st[s].rho = tot[0];
    B.st[s].rho_old = tot[0];
}

// ---- K1 ------------------------------------------------------------------------------------
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k1(BArgs B, int t0, const T* __restrict__ p, T* __restrict__ q, T* __restrict__ r, const double* __restrict__ b,
       const double* __restrict__ x64) {
    using Acc = T;
    __shared__ double prod_d[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const BState* S0 = &B.st[s];
    if (S0->halt) return;
    const int flag = S0->flag;
    if (flag != BS_NORMAL && flag != BS_RESID) return;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[2] = {0.0, 0.0};
    if (flag == BS_NORMAL) {
        tile_spmv<Acc>(r0, r1, B.rp, B.ci, Prec<T>::val(B), p, reinterpret_cast<Acc*>(prod_d), [&](int i, Acc sum) {
            const T qi = (T)sum;
            q[i] = qi;
            red[0] += (double)p[i] * (double)qi;
        });
    } else {
        const T* dinv = Prec<T>::dinv(B);
        tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x64, prod_d, [&](int i, double sum) {
            const double rh = b[i] - sum;
            const T ri = (T)rh;
            r[i] = ri;
            red[0] += rh * rh;
            red[1] += (double)ri * (double)(T)(dinv[i] * ri);
        });
    }
    double tot[2];
    if (!sys_sum<2>(red, B, tile, s, tot)) return;
    BState& S = B.st[s];
    if (flag == BS_NORMAL) {
        const double pq = tot[0];
        if (fabs(pq) < kEps * 1e6 * fabs(S.rho)) {  // cg_solver.cpp:109-120: restart
            S.iter += 1;
            S.restarts += 1;
            S.flag = BS_FLUSH;
            S.reason = BR_RESTART;
            if (S.iter >= B.max_iter) S.halt = 1;
        } else {
            S.alpha = S.rho / pq;
        }
        return;
    }
    // RESID: the FP64 true residual replaced r
    const double rn = sqrt(tot[0]);
    S.true_rel = rn / S.bnorm;
    S.rmax = rn;
    if (S.reason != BR_RESTART) S.reliable += 1;
    if (S.reason == BR_CHECK && S.true_rel < B.tol) {
        S.rel = S.true_rel;
        S.done = S.conv = S.halt = 1;
        S.flag = BS_NORMAL;
        return;
    }
    if (S.reason == BR_RESTART) {
        S.rho = tot[1];
        S.rho_old = tot[1];
        S.beta = 0.0;
    } else {
        S.rho_old = S.rho;
        S.rho = tot[1];
        S.beta = S.rho / S.rho_old;
    }
    S.flag = BS_PUPDATE;
}

// ---- K2 ------------------------------------------------------------------------------------
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k2(BArgs B, int t0, const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ r,
       double* __restrict__ x64, float* __restrict__ xlo) {
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const BState* S0 = &B.st[s];
    if (S0->halt || S0->flag != BS_NORMAL) return;
    const double alpha = S0->alpha;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    const T* dinv = Prec<T>::dinv(B);
    double red[2] = {0.0, 0.0};
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
        if (MIXED) xlo[i] = (float)((double)xlo[i] + alpha * (double)p[i]);
        else x64[i] = x64[i] + alpha * (double)p[i];
        const T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        const double rd = (double)ri;
        red[0] += rd * rd;
        red[1] += rd * (double)(T)(dinv[i] * ri);
    }
    double tot[2];
    if (!sys_sum<2>(red, B, tile, s, tot)) return;
    BState& S = B.st[s];
    const double rn = sqrt(tot[0]);
    S.rel = rn / S.bnorm;
    S.iter += 1;
    if (S.rel < B.tol) {  // cg_solver.cpp:143
        if (MIXED) {
            S.flag = BS_FLUSH;
            S.reason = BR_CHECK;
        } else {
            S.done = S.conv = S.halt = 1;
        }
        return;
    }
    if (MIXED) {
        S.rmax = fmax(S.rmax, rn);
        if (rn < B.delta * S.rmax) {
            S.flag = BS_FLUSH;
            S.reason = BR_RELIABLE;
            if (S.iter >= B.max_iter) S.halt = 1;
            return;
        }
    }
    S.rho_old = S.rho;
    S.rho = tot[1];
    if (fabs(S.rho_old) < kEps * 1e8 * fabs(S.rho)) {  // cg_solver.cpp:157-167: restart
        S.restarts += 1;
        S.flag = BS_FLUSH;
        S.reason = BR_RESTART;
        if (S.iter >= B.max_iter) S.halt = 1;
        return;
    }
    S.beta = S.rho / S.rho_old;
    if (S.iter >= B.max_iter) S.halt = 1;
}

// ---- K3 ------------------------------------------------------------------------------------
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k3(BArgs B, int t0, T* __restrict__ p, const T* __restrict__ r, double* __restrict__ x64,
       float* __restrict__ xlo) {
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const BState* S0 = &B.st[s];
    if (S0->halt) return;
    const int flag = S0->flag;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    if (flag == BS_NORMAL || flag == BS_PUPDATE) {
        const double beta = S0->beta;
        const T* dinv = Prec<T>::dinv(B);
        for (int i = r0 + threadIdx.x; i < r1; i += kBlock)
            p[i] = (T)((double)(T)(dinv[i] * r[i]) + beta * (double)p[i]);
        if (flag == BS_PUPDATE && sys_ticket(B, s)) B.st[s].flag = BS_NORMAL;
    } else if (flag == BS_FLUSH) {
        if (MIXED)
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        if (sys_ticket(B, s)) B.st[s].flag = BS_RESID;
    }
}

// ---- final: x64 += x_lo (all rows) and the FP64 true residual per system -------------------
__global__ void k_b_flush_all(int n, double* __restrict__ x64, float* __restrict__ xlo) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x64[i] += (double)xlo[i];
        xlo[i] = 0.0f;
    }
}

__global__ void __launch_bounds__(kBlock)
k_b_final(BArgs B, int t0, const double* __restrict__ b, const double* __restrict__ x) {
    __shared__ double prod[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[1] = {0.0};
    tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x, prod, [&](int i, double sum) {
        const double ri = b[i] - sum;
        red[0] += ri * ri;
    });
    double tot[1];
    if (!sys_sum<1>(red, B, tile, s, tot)) return;
    BState& S = B.st[s];
    S.true_rel = S.bnorm > kEps ? sqrt(tot[0]) / S.bnorm : sqrt(tot[0]);
}

// =============================================================================
// host side
// =============================================================================
struct Batch {
    Ctx* ctx = nullptr;
    int S = 0;  // systems
    int N = 0;  // rows
    long long nnz = 0;
    std::vector<int> h_row_off;  // user order, S+1
    std::vector<int> h_rp;       // global row_ptr, N+1
    std::vector<int> order;      // internal position -> user system id (sorted by nnz, descending)
    DevBuf rp, ci, v64, v32, dinv64, dinv32;
    // schedule (depends on tile_rows)
    int tile_rows = 0;
    int T = 0;
    DevBuf tile_r0, tile_r1, tile_sys, sys_t0, sys_t1;
    std::vector<int> h_sys_t0, h_sys_t1;
    std::vector<std::pair<int, int>> waves;  // internal system ranges
    // per-solve work (reused)
    DevBuf st, cnt, partials, r64, r, p, q, xlo;
    int work_prec = -1;
    // cached graphs per (wave, precision)
    struct G { hipGraphExec_t exec = nullptr; int wave = -1, prec = -1, chunk = 0; };
    std::vector<G> graphs;
    ~Batch() { drop_graphs(); }
    void drop_graphs() {
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
};

static void build_tiles(Batch& B, int tile_rows) {
    std::vector<int> r0, r1, sys;
    B.h_sys_t0.assign(B.S, 0);
    B.h_sys_t1.assign(B.S, 0);
    for (int k = 0; k < B.S; ++k) {
        const int s = B.order[k];
        const int a = B.h_row_off[s], e = B.h_row_off[s + 1];
        B.h_sys_t0[k] = (int)r0.size();
        int i = a;
        while (i < e) {
            int j = i;
            const int base = B.h_rp[i];
            if (B.h_rp[i + 1] - base > kLdsNnz) {
                j = i + 1;
            } else {
                while (j < e && j - i < tile_rows && B.h_rp[j + 1] - base <= kLdsNnz) ++j;
            }
            r0.push_back(i);
            r1.push_back(j);
            sys.push_back(k);
            i = j;
        }
        if (a == e) {  // empty system: one empty tile keeps the reduction well-formed
            r0.push_back(a);
            r1.push_back(a);
            sys.push_back(k);
        }
        B.h_sys_t1[k] = (int)r0.size();
    }
    B.T = (int)r0.size();
    auto up = [&](DevBuf& d, const std::vector<int>& h) {
        d = DevBuf(B.ctx, std::max<size_t>(h.size(), 1) * sizeof(int));
        if (!h.empty()) HSPS_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    up(B.tile_r0, r0);
    up(B.tile_r1, r1);
    up(B.tile_sys, sys);
    up(B.sys_t0, B.h_sys_t0);
    up(B.sys_t1, B.h_sys_t1);
    B.partials = DevBuf(B.ctx, sizeof(double) * 2 * (size_t)std::max(B.T, 1));
    B.tile_rows = tile_rows;
    B.drop_graphs();
}

// Waves: consecutive systems (largest first) whose FP32 working set
// (8 B/nnz + ~36 B/row of vectors) stays under the MALL budget.
static void build_waves(Batch& B, size_t budget) {
    B.waves.clear();
    int start = 0;
    size_t acc = 0;
    for (int k = 0; k < B.S; ++k) {
        const int s = B.order[k];
        const size_t rows = B.h_row_off[s + 1] - B.h_row_off[s];
        const size_t nz = B.h_rp[B.h_row_off[s + 1]] - B.h_rp[B.h_row_off[s]];
        const size_t ws = nz * 8 + rows * 36;
        if (k > start && acc + ws > budget) {
            B.waves.push_back({start, k});
            start = k;
            acc = 0;
        }
        acc += ws;
    }
    if (start < B.S) B.waves.push_back({start, B.S});
}

static BArgs args_of(const Batch& B, const hsps_params& p) {
    BArgs a;
    a.rp = B.rp.as<int>();
    a.ci = B.ci.as<int>();
    a.v64 = B.v64.as<double>();
    a.v32 = B.v32.as<float>();
    a.dinv64 = B.dinv64.as<double>();
    a.dinv32 = B.dinv32.as<float>();
    a.tile_r0 = B.tile_r0.as<int>();
    a.tile_r1 = B.tile_r1.as<int>();
    a.tile_sys = B.tile_sys.as<int>();
    a.sys_t0 = B.sys_t0.as<int>();
    a.sys_t1 = B.sys_t1.as<int>();
    a.st = B.st.as<BState>();
    a.cnt = B.cnt.as<unsigned>();
    a.partials = B.partials.as<double>();
    a.T = B.T;
    a.tol = p.tol;
    a.delta = p.reliable_delta > 0 ? p.reliable_delta : 0.1;
    a.max_iter = p.max_iter;
    return a;
}

template <class T, bool MIXED>
static void enqueue_iters(const Batch& B, const BArgs& a, int t0, int nt, int iters, const double* b, double* x,
                          hipStream_t s) {
    T* r = B.r.as<T>();
    T* p = B.p.as<T>();
    T* q = B.q.as<T>();
    float* xl = B.xlo.as<float>();
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL((k_b_k1<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, (const T*)p, q, r, b, (const double*)x);
        hipLaunchKernelGGL((k_b_k2<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, (const T*)p, (const T*)q, r, x, xl);
        hipLaunchKernelGGL((k_b_k3<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, p, (const T*)r, x, xl);
    }
    HSPS_CHECK_LAUNCH();
}

static void ensure_work(Batch& B, int prec) {
    const size_t n = (size_t)std::max(B.N, 1);
    if (!B.st.p) {
        B.st = DevBuf(B.ctx, sizeof(BState) * (size_t)std::max(B.S, 1));
        B.cnt = DevBuf(B.ctx, sizeof(unsigned) * (size_t)std::max(B.S, 1));
        HSPS_HIP(hipMemset(B.cnt.p, 0, B.cnt.bytes));
        B.r64 = DevBuf(B.ctx, n * sizeof(double));
    }
    if (B.work_prec != prec) {
        const size_t vb = prec == HSPS_MIXED ? 4 : 8;
        B.r = DevBuf(B.ctx, n * vb);
        B.p = DevBuf(B.ctx, n * vb);
        B.q = DevBuf(B.ctx, n * vb);
        B.xlo = DevBuf(B.ctx, prec == HSPS_MIXED ? n * sizeof(float) : 0);
        B.work_prec = prec;
        B.drop_graphs();
    }
}

static void launch_init(Batch& B, const BArgs& a, const double* b, double* x, int prec, hipStream_t s) {
    if (B.T == 0) return;
    hipLaunchKernelGGL(k_b_init1, dim3(B.T), dim3(kBlock), 0, s, a, 0, b, (const double*)x, B.r64.as<double>());
    if (prec == HSPS_MIXED)
        hipLaunchKernelGGL((k_b_init2<float, true>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                           (const double*)B.r64.as<double>(), B.r.as<float>(), B.p.as<float>(), B.xlo.as<float>());
    else
        hipLaunchKernelGGL((k_b_init2<double, false>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                           (const double*)B.r64.as<double>(), B.r.as<double>(), B.p.as<double>(), (float*)nullptr);
    HSPS_CHECK_LAUNCH();
}

static void launch_chunk(Batch& B, const BArgs& a, int w, int chunk, const double* b, double* x, int prec,
                         hipStream_t s, bool use_graph) {
    const int k0 = B.waves[w].first, k1 = B.waves[w].second;
    const int t0 = B.h_sys_t0[k0], nt = B.h_sys_t1[k1 - 1] - t0;
    if (nt <= 0) return;
    auto direct = [&]() {
        if (prec == HSPS_MIXED) enqueue_iters<float, true>(B, a, t0, nt, chunk, b, x, s);
        else enqueue_iters<double, false>(B, a, t0, nt, chunk, b, x, s);
    };
    if (!use_graph) { direct(); return; }
    Batch::G* g = nullptr;
    for (auto& c : B.graphs)
        if (c.wave == w && c.prec == prec && c.chunk == chunk) g = &c;
    if (!g) {
        hipGraph_t graph = nullptr;
        HSPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        direct();
        HSPS_HIP(hipStreamEndCapture(s, &graph));
        Batch::G ng;
        ng.wave = w;
        ng.prec = prec;
        ng.chunk = chunk;
        HSPS_HIP(hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        B.graphs.push_back(ng);
        g = &B.graphs.back();
    }
    HSPS_HIP(hipGraphLaunch(g->exec, s));
}

// Runs the solve; returns per-system states in internal order.
static std::vector<BState> run_batch(Batch& B, const double* b, double* x, const hsps_params& p, bool use_graph = true,
                                     int max_chunks = -1) {
    hipStream_t s = B.ctx->stream;
    const int prec = p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64;
    ensure_work(B, prec);
    const BArgs a = args_of(B, p);
    launch_init(B, a, b, x, prec, s);
    const int chunk = std::max(1, std::min(p.check_every > 0 ? p.check_every : 32, 64));
    std::vector<BState> host(B.S);
    BState* pin = nullptr;
    HSPS_HIP(hipHostMalloc((void**)&pin, sizeof(BState) * (size_t)std::max(B.S, 1) * 2, hipHostMallocDefault));
    hipEvent_t ev[2];
    HSPS_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HSPS_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    try {
        for (int w = 0; w < (int)B.waves.size(); ++w) {
            const int k0 = B.waves[w].first, k1 = B.waves[w].second;
            const int ns = k1 - k0;
            int k = 0, chunks = 0;
            bool pending = false;
            const int cap = (p.max_iter + chunk - 1) / chunk + 2;
            for (;;) {
                launch_chunk(B, a, w, chunk, b, x, prec, s, use_graph);
                ++chunks;
                BState* slot = pin + (size_t)(k & 1) * B.S;
                HSPS_HIP(hipMemcpyAsync(slot + k0, B.st.as<BState>() + k0, sizeof(BState) * ns, hipMemcpyDeviceToHost, s));
                HSPS_HIP(hipEventRecord(ev[k & 1], s));
                bool all = false;
                if (pending) {
                    HSPS_HIP(hipEventSynchronize(ev[(k - 1) & 1]));
                    const BState* prev = pin + (size_t)((k - 1) & 1) * B.S;
                    all = true;
                    for (int i = k0; i < k1 && all; ++i) all = prev[i].halt != 0;
                }
                pending = true;
                ++k;
                if (all || chunks >= cap || (max_chunks > 0 && chunks >= max_chunks)) break;
            }
        }
        if (prec == HSPS_MIXED)
            hipLaunchKernelGGL(k_b_flush_all, dim3(grid_for(B.N)), dim3(kBlock), 0, s, B.N, x, B.xlo.as<float>());
        if (B.T) hipLaunchKernelGGL(k_b_final, dim3(B.T), dim3(kBlock), 0, s, a, 0, b, (const double*)x);
        HSPS_CHECK_LAUNCH();
        HSPS_HIP(hipMemcpyAsync(pin, B.st.p, sizeof(BState) * B.S, hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
        std::memcpy(host.data(), pin, sizeof(BState) * B.S);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipHostFree(pin);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipHostFree(pin);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return host;
}

// Project autotuner: times a few chunks of the first wave for each tile size
// on scratch copies of x, keeps the fastest.
static void autotune(Batch& B, const double* b, const double* x0, const hsps_params& p) {
    const int cands[] = {64, 128, 256};
    hipStream_t s = B.ctx->stream;
    DevBuf xs(B.ctx, (size_t)std::max(B.N, 1) * sizeof(double));
    hsps_params q = p;
    q.check_every = 16;
    int best = 256;
    double best_t = 1e30;
    for (int tr : cands) {
        build_tiles(B, tr);
        HSPS_HIP(hipMemcpyAsync(xs.p, x0, (size_t)B.N * sizeof(double), hipMemcpyDeviceToDevice, s));
        (void)run_batch(B, b, xs.as<double>(), q, false, 1);  // warm
        HSPS_HIP(hipMemcpyAsync(xs.p, x0, (size_t)B.N * sizeof(double), hipMemcpyDeviceToDevice, s));
        HSPS_HIP(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        (void)run_batch(B, b, xs.as<double>(), q, false, 2);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt < best_t) { best_t = dt; best = tr; }
    }
    build_tiles(B, best);
}

}  // namespace hsps_dev

using namespace hsps_dev;

struct hsps_batch_s : Batch {};

extern "C" {

int hsps_batch_create(hsps_ctx ctx, int S, const int* row_off, const int* rp, const int* ci, const double* v,
                      hsps_batch* out) {
    return guarded([&] {
        if (!ctx || !out || S < 0 || !row_off || !rp) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: bad argument");
        const int N = row_off[S];
        if (row_off[0] != 0 || N < 0) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: row_off[0] must be 0");
        const long long nnz = rp[N];
        if (rp[0] != 0 || nnz < 0 || (nnz > 0 && (!ci || !v))) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: bad row_ptr");
        std::vector<double> d64(N);
        for (int s = 0; s < S; ++s) {
            if (row_off[s + 1] < row_off[s]) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: row_off not monotone");
            for (int i = row_off[s]; i < row_off[s + 1]; ++i) {
                if (rp[i + 1] < rp[i]) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: row_ptr not monotone");
                double d = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    if (ci[k] < row_off[s] || ci[k] >= row_off[s + 1])
                        throw Error(HSPS_ERR_INVALID, "hsps_batch_create: column outside its system's block (row " +
                                                          std::to_string(i) + ")");
                    if (ci[k] == i) d = v[k];
                }
                if (std::abs(d) < kEps)  // jacobi.cpp:28-30
                    throw Error(HSPS_ERR_ZERO_DIAG, "JacobiPreconditioner: zero diagonal at row " + std::to_string(i));
                d64[i] = 1.0 / d;
            }
        }
        auto* B = new hsps_batch_s;
        try {
            B->ctx = ctx;
            B->S = S;
            B->N = N;
            B->nnz = nnz;
            B->h_row_off.assign(row_off, row_off + S + 1);
            B->h_rp.assign(rp, rp + N + 1);
            ctx->set_current();
            B->rp = DevBuf(ctx, sizeof(int) * (size_t)(N + 1));
            B->ci = DevBuf(ctx, sizeof(int) * (size_t)std::max<long long>(nnz, 1));
            B->v64 = DevBuf(ctx, sizeof(double) * (size_t)std::max<long long>(nnz, 1));
            B->v32 = DevBuf(ctx, sizeof(float) * (size_t)std::max<long long>(nnz, 1));
            B->dinv64 = DevBuf(ctx, sizeof(double) * (size_t)std::max(N, 1));
            B->dinv32 = DevBuf(ctx, sizeof(float) * (size_t)std::max(N, 1));
            HSPS_HIP(hipMemcpy(B->rp.p, rp, sizeof(int) * (size_t)(N + 1), hipMemcpyHostToDevice));
            if (nnz) {
                HSPS_HIP(hipMemcpy(B->ci.p, ci, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
                HSPS_HIP(hipMemcpy(B->v64.p, v, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
                std::vector<float> f(v, v + nnz);
                HSPS_HIP(hipMemcpy(B->v32.p, f.data(), sizeof(float) * (size_t)nnz, hipMemcpyHostToDevice));
            }
            if (N) {
                HSPS_HIP(hipMemcpy(B->dinv64.p, d64.data(), sizeof(double) * (size_t)N, hipMemcpyHostToDevice));
                std::vector<float> f(d64.begin(), d64.end());
                HSPS_HIP(hipMemcpy(B->dinv32.p, f.data(), sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
            }
            B->order.resize(S);
            std::iota(B->order.begin(), B->order.end(), 0);
            std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b) {
                return rp[row_off[a + 1]] - rp[row_off[a]] > rp[row_off[b + 1]] - rp[row_off[b]];
            });
            build_tiles(*B, 256);
            B->tile_rows = 0;  // not tuned yet
            build_waves(*B, (size_t)192 << 20);
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
        return HSPS_OK;
    });
}

int hsps_batch_destroy(hsps_batch B) {
    return guarded([&] {
        if (B) {
            (void)hipSetDevice(B->ctx->device);
            (void)hipStreamSynchronize(B->ctx->stream);
            delete B;
        }
        return HSPS_OK;
    });
}

int hsps_batch_solve(hsps_batch B, const double* b, double* x, const hsps_params* pp, int* iters, int* conv,
                     double* true_res) {
    return guarded([&] {
        if (!B || (B->N > 0 && (!b || !x))) throw Error(HSPS_ERR_INVALID, "hsps_batch_solve: null argument");
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        B->ctx->set_current();
        if (p.batch_tile_rows > 0) {
            if (p.batch_tile_rows != B->tile_rows) build_tiles(*B, std::min(p.batch_tile_rows, kMaxBlockRows));
        } else if (B->tile_rows == 0) {
            ensure_work(*B, p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64);
            autotune(*B, b, x, p);
        }
        std::vector<BState> st = run_batch(*B, b, x, p);
        int n_conv = 0;
        for (int k = 0; k < B->S; ++k) {
            const int s = B->order[k];
            const BState& S = st[k];
            const int c = S.conv;
            if (iters) iters[s] = S.iter;
            if (conv) conv[s] = c;
            if (true_res) true_res[s] = S.true_rel;
            n_conv += c;
        }
        return n_conv == B->S ? 1 : 0;
    });
}

}  // extern "C"
