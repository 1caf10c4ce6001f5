// =============================================================================
// batch.hip — batched PCG + Jacobi for many independent systems (no reference
// counterpart; the reference solves one system per call, cg_solver.cpp:37-191
// gives the per-system algorithm followed here).
//
// Layout: the systems' CSRs are concatenated (block-diagonal, global column
// indices).  Systems are sorted by nonzeros (largest first) and split in two
// classes:
//  * small systems (rows <= kSmallRows): one persistent 1024-thread
//    workgroup per system runs the whole PCG loop with workgroup barriers
//    only — no launches per iteration, vectors stay in L2;
//  * large systems: cut into tiles of <= tile_rows rows (one workgroup per
//    tile, tile_rows autotuned), grouped into waves by a working-set budget
//    (default: one wave).  Per iteration five launches over the wave's tiles:
//      K1  q = A p, p.q partials  (FLUSH state: FP64 true residual r = b - A x)
//      F1  per-system finish: alpha (or breakdown / reliable-update decisions)
//      K2  x += alpha p, r -= alpha q, r.r / r.z partials
//      F2  per-system finish: rel, rho, beta, convergence
//      K3  p = D^-1 r + beta p   (FLUSH state: x64 += x_lo)
//    The chunk of iterations is captured once per wave in a hipGraph and
//    replayed; the host polls the wave's states one chunk behind.
// The two classes run concurrently on two streams.
// Mixed precision: FP32 matrix values and Krylov vectors, FP64 reductions and
// scalars, x accumulated in FP64 through an FP32 low-order part flushed at
// every reliable update (|r| < delta * max|r| since the last one, or |r| <
// tol to confirm convergence against the FP64 true residual).
// =============================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>

#include "device_common.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr int kSmallRows = 20000;   // persistent one-workgroup path up to this size
constexpr int kSmallThreads = 1024;

enum : int { BS_NORMAL = 0, BS_FLUSH = 1, BS_PUPDATE = 2 };
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
    const int* sys_row0;  // internal order
    const int* sys_row1;
    BState* st;
    double* partials;
    int T;  // total tiles
    double tol, delta;
    int max_iter;
};

template <class T> struct Prec;
template <> struct Prec<double> {
    static __device__ const double* val(const BArgs& B) { return B.v64; }
    static __device__ const double* dinv(const BArgs& B) { return B.dinv64; }
};
template <> struct Prec<float> {
    static __device__ const float* val(const BArgs& B) { return B.v32; }
    static __device__ const float* dinv(const BArgs& B) { return B.dinv32; }
};

// Row sums of a tile (rows r0..r1) of A*x, LDS-staged; f(row, sum) once per
// row from one lane.  Same staging rule as the single-system SpMV.
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

// Per-system finish: one workgroup per system of [k0, k0+gridDim.x); sums the
// system's tile partials in tile order and runs op(S, tot) in thread 0.
template <int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_b_finish(BArgs B, int k0, Op op) {
    const int s = k0 + blockIdx.x;
    BState* S = &B.st[s];
    if (!op.active(*S)) return;
    const int t0 = B.sys_t0[s], t1 = B.sys_t1[s];
    __shared__ double red[NV * (kBlock / 64)];
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += B.partials[(size_t)k * B.T + t];
    }
    block_sum<NV>(acc, red);
    if (threadIdx.x == 0) op(*S, acc, B);
}

template <int NV>
__device__ __forceinline__ void tile_partial(double (&v)[NV], const BArgs& B, int tile) {
    __shared__ double red[NV * (kBlock / 64)];
    block_sum<NV>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) B.partials[(size_t)k * B.T + tile] = v[k];
    }
}

// ---- init: r64 = b - A x ; |r|, |b| --------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_b_init1(BArgs B, int t0, const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r64) {
    __shared__ double prod[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[2] = {0.0, 0.0};
    tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x, prod, [&](int i, double sum) {
        const double ri = b[i] - sum;
        r64[i] = ri;
        red[0] += ri * ri;
    });
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) red[1] += b[i] * b[i];
    tile_partial<2>(red, B, tile);
}

struct OpInit1 {
    __device__ bool active(const BState&) const { return true; }
    __device__ void operator()(BState& S0, const double* tot, const BArgs& B) const {
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
        S0 = S;
    }
};

// r = (T) r64, x_lo = 0, p = D^-1 r, r.z partials
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
    tile_partial<1>(red, B, tile);
}

struct OpInit2 {
    __device__ bool active(const BState& S) const { return !S.halt; }
    __device__ void operator()(BState& S, const double* tot, const BArgs&) const {
        S.rho = tot[0];
        S.rho_old = tot[0];
    }
};

// ---- K1 / F1 -------------------------------------------------------------------------------
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k1(BArgs B, int t0, const T* __restrict__ p, T* __restrict__ q, T* __restrict__ r, const double* __restrict__ b,
       const double* __restrict__ x64) {
    __shared__ double prod_d[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int s = B.tile_sys[tile];
    const BState* S0 = &B.st[s];
    if (S0->halt) return;
    const int flag = S0->flag;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[2] = {0.0, 0.0};
    if (flag != BS_FLUSH) {
        tile_spmv<T>(r0, r1, B.rp, B.ci, Prec<T>::val(B), p, reinterpret_cast<T*>(prod_d), [&](int i, T sum) {
            q[i] = sum;
            red[0] += (double)p[i] * (double)sum;
        });
    } else {  // x was flushed by the previous K3: FP64 true residual replaces r
        const T* dinv = Prec<T>::dinv(B);
        tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x64, prod_d, [&](int i, double sum) {
            const double rh = b[i] - sum;
            const T ri = (T)rh;
            r[i] = ri;
            red[0] += rh * rh;
            red[1] += (double)ri * (double)(T)(dinv[i] * ri);
        });
    }
    tile_partial<2>(red, B, tile);
}

struct OpF1 {
    __device__ bool active(const BState& S) const { return !S.halt; }
    __device__ void operator()(BState& S, const double* tot, const BArgs& B) const {
        if (S.flag != BS_FLUSH) {
            const double pq = tot[0];
            if (fabs(pq) < kEps * 1e6 * fabs(S.rho)) {  // cg_solver.cpp:109-120: restart
                S.iter += 1;
                S.restarts += 1;
                S.flag = BS_FLUSH;
                S.reason = BR_RESTART;
                if (S.iter >= B.max_iter) S.halt = 1;
            } else {
                S.alpha = S.rho / pq;
                S.flag = BS_NORMAL;
            }
            return;
        }
        // FLUSH: the FP64 true residual replaced r
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
        S.reason = BR_NONE;
    }
};

// ---- K2 / F2 -------------------------------------------------------------------------------
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
    tile_partial<2>(red, B, tile);
}

template <bool MIXED>
struct OpF2 {
    __device__ bool active(const BState& S) const { return !S.halt && S.flag == BS_NORMAL; }
    __device__ void operator()(BState& S, const double* tot, const BArgs& B) const {
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
};

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
    if (flag == BS_FLUSH) {
        if (MIXED)
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        return;
    }
    const double beta = S0->beta;
    const T* dinv = Prec<T>::dinv(B);
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock)
        p[i] = (T)((double)(T)(dinv[i] * r[i]) + beta * (double)p[i]);
}

// ---- final: x64 += x_lo and the FP64 true residual per (large) system ----------------------
__global__ void k_b_flush_range(int a, int e, double* __restrict__ x64, float* __restrict__ xlo) {
    for (int i = a + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        x64[i] += (double)xlo[i];
        xlo[i] = 0.0f;
    }
}

__global__ void __launch_bounds__(kBlock)
k_b_final(BArgs B, int t0, const double* __restrict__ b, const double* __restrict__ x) {
    __shared__ double prod[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[1] = {0.0};
    tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, x, prod, [&](int i, double sum) {
        const double ri = b[i] - sum;
        red[0] += ri * ri;
    });
    tile_partial<1>(red, B, tile);
}

struct OpFinal {
    __device__ bool active(const BState&) const { return true; }
    __device__ void operator()(BState& S, const double* tot, const BArgs&) const {
        S.true_rel = S.bnorm > kEps ? sqrt(tot[0]) / S.bnorm : sqrt(tot[0]);
    }
};

// ---- small systems: one persistent workgroup per system ---------------------------------------
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double* red) {
    constexpr int W = kSmallThreads / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();  // red[] reuse across calls
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * W + wid] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) s += red[k * W + w];
        v[k] = s;  // every thread gets the total
    }
}

// The whole PCG of cg_solver.cpp:37-191 for one system, with the same
// reliable-update refinement as the tiled path in mixed mode.
template <class T, bool MIXED>
__global__ void __launch_bounds__(kSmallThreads)
k_b_small(BArgs B, int k0, const double* __restrict__ b, double* __restrict__ x64, T* __restrict__ r,
          T* __restrict__ p, T* __restrict__ q, float* __restrict__ xlo) {
    __shared__ double red[4 * (kSmallThreads / 64)];
    const int s = k0 + blockIdx.x;
    const int a = B.sys_row0[s], e = B.sys_row1[s];
    const int tid = threadIdx.x;
    const int* __restrict__ rp = B.rp;
    const int* __restrict__ ci = B.ci;
    const T* __restrict__ val = Prec<T>::val(B);
    const T* __restrict__ dinv = Prec<T>::dinv(B);
    // true residual into r (rounded), returns (|r64|^2, r.z)
    auto true_residual = [&](double (&o)[2]) {
        o[0] = o[1] = 0.0;
        for (int i = a + tid; i < e; i += kSmallThreads) {
            double sum = 0.0;
            for (int k = rp[i]; k < rp[i + 1]; ++k) sum += B.v64[k] * x64[ci[k]];
            const double rh = b[i] - sum;
            const T ri = (T)rh;
            r[i] = ri;
            o[0] += rh * rh;
            o[1] += (double)ri * (double)(T)(dinv[i] * ri);
        }
        wg_sum<2>(o, red);
    };
    BState S{};
    double bb[1] = {0.0};
    for (int i = a + tid; i < e; i += kSmallThreads) bb[0] += b[i] * b[i];
    wg_sum<1>(bb, red);
    S.bnorm = sqrt(bb[0]);
    double o[2];
    true_residual(o);
    const double r0n = sqrt(o[0]);
    S.rmax = r0n;
    S.rel = S.bnorm < kEps ? r0n : r0n / S.bnorm;
    S.true_rel = S.rel;
    if (S.bnorm < kEps || S.rel < B.tol) {
        S.done = S.conv = S.halt = 1;
    } else {
        // p = z = D^-1 r ; rho = r.z
        for (int i = a + tid; i < e; i += kSmallThreads) {
            p[i] = (T)(dinv[i] * r[i]);
            if (MIXED) xlo[i] = 0.0f;
        }
        S.rho = S.rho_old = o[1];
        __syncthreads();
        while (S.iter < B.max_iter) {
            // q = A p ; p.q
            double v1[1] = {0.0};
            for (int i = a + tid; i < e; i += kSmallThreads) {
                T sum = 0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) sum += val[k] * p[ci[k]];
                q[i] = sum;
                v1[0] += (double)p[i] * (double)sum;
            }
            wg_sum<1>(v1, red);
            int reason = BR_NONE;
            const double pq = v1[0];
            if (fabs(pq) < kEps * 1e6 * fabs(S.rho)) {
                S.iter += 1;
                S.restarts += 1;
                reason = BR_RESTART;
            } else {
                const double alpha = S.rho / pq;
                double v2[2] = {0.0, 0.0};
                for (int i = a + tid; i < e; i += kSmallThreads) {
                    if (MIXED) xlo[i] = (float)((double)xlo[i] + alpha * (double)p[i]);
                    else x64[i] = x64[i] + alpha * (double)p[i];
                    const T ri = (T)((double)r[i] - alpha * (double)q[i]);
                    r[i] = ri;
                    v2[0] += (double)ri * (double)ri;
                    v2[1] += (double)ri * (double)(T)(dinv[i] * ri);
                }
                wg_sum<2>(v2, red);
                const double rn = sqrt(v2[0]);
                S.rel = rn / S.bnorm;
                S.iter += 1;
                if (MIXED) S.rmax = fmax(S.rmax, rn);
                if (S.rel < B.tol) {
                    if (!MIXED) { S.done = S.conv = 1; break; }
                    reason = BR_CHECK;
                } else if (MIXED && rn < B.delta * S.rmax) {
                    reason = BR_RELIABLE;
                } else {
                    S.rho_old = S.rho;
                    S.rho = v2[1];
                    if (fabs(S.rho_old) < kEps * 1e8 * fabs(S.rho)) {
                        S.restarts += 1;
                        reason = BR_RESTART;
                    } else {
                        S.beta = S.rho / S.rho_old;
                    }
                }
            }
            if (reason != BR_NONE) {
                if (MIXED) {
                    for (int i = a + tid; i < e; i += kSmallThreads) {
                        x64[i] += (double)xlo[i];
                        xlo[i] = 0.0f;
                    }
                }
                __syncthreads();
                true_residual(o);
                const double rn = sqrt(o[0]);
                S.true_rel = rn / S.bnorm;
                S.rmax = rn;
                if (reason != BR_RESTART) S.reliable += 1;
                if (reason == BR_CHECK && S.true_rel < B.tol) {
                    S.rel = S.true_rel;
                    S.done = S.conv = 1;
                    break;
                }
                if (reason == BR_RESTART) {
                    S.rho = S.rho_old = o[1];
                    S.beta = 0.0;
                } else {
                    S.rho_old = S.rho;
                    S.rho = o[1];
                    S.beta = S.rho / S.rho_old;
                }
            }
            if (S.iter >= B.max_iter) break;
            for (int i = a + tid; i < e; i += kSmallThreads)
                p[i] = (T)((double)(T)(dinv[i] * r[i]) + S.beta * (double)p[i]);
            __syncthreads();
        }
        S.halt = 1;
        if (MIXED)
            for (int i = a + tid; i < e; i += kSmallThreads) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        __syncthreads();
        double t[1] = {0.0};
        for (int i = a + tid; i < e; i += kSmallThreads) {
            double sum = 0.0;
            for (int k = rp[i]; k < rp[i + 1]; ++k) sum += B.v64[k] * x64[ci[k]];
            const double rh = b[i] - sum;
            t[0] += rh * rh;
        }
        wg_sum<1>(t, red);
        S.true_rel = sqrt(t[0]) / S.bnorm;
    }
    if (tid == 0) B.st[s] = S;
}

// =============================================================================
// host side
// =============================================================================
struct Batch {
    Ctx* ctx = nullptr;
    hipStream_t aux = nullptr;  // small-system stream
    hipEvent_t join = nullptr, fork = nullptr;
    int S = 0;  // systems
    int N = 0;  // rows
    long long nnz = 0;
    int n_large = 0;  // internal systems [0, n_large) are tiled; the rest are small
    std::vector<int> h_row_off;  // user order, S+1
    std::vector<int> h_rp;       // global row_ptr, N+1
    std::vector<int> order;      // internal position -> user system id (by nnz, descending)
    DevBuf rp, ci, v64, v32, dinv64, dinv32, sys_row0, sys_row1;
    // tiled schedule (depends on tile_rows)
    int tile_rows = 0;
    int T = 0;
    DevBuf tile_r0, tile_r1, tile_sys, sys_t0, sys_t1;
    std::vector<int> h_sys_t0, h_sys_t1;
    std::vector<std::pair<int, int>> waves;  // internal system ranges within [0, n_large)
    // per-solve work (reused)
    DevBuf st, partials, r64, r, p, q, xlo;
    int work_prec = -1;
    struct G { hipGraphExec_t exec = nullptr; int wave = -1, prec = -1, chunk = 0; };
    std::vector<G> graphs;
    ~Batch() {
        drop_graphs();
        if (join) (void)hipEventDestroy(join);
        if (fork) (void)hipEventDestroy(fork);
        if (aux) (void)hipStreamDestroy(aux);
    }
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
    for (int k = 0; k < B.n_large; ++k) {
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

// Waves over the large systems: consecutive systems whose FP32 working set
// (8 B/nnz + ~36 B/row of vectors) stays under the MALL budget.
static void build_waves(Batch& B, size_t budget) {
    B.waves.clear();
    int start = 0;
    size_t acc = 0;
    for (int k = 0; k < B.n_large; ++k) {
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
    if (start < B.n_large) B.waves.push_back({start, B.n_large});
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
    a.sys_row0 = B.sys_row0.as<int>();
    a.sys_row1 = B.sys_row1.as<int>();
    a.st = B.st.as<BState>();
    a.partials = B.partials.as<double>();
    a.T = B.T;
    a.tol = p.tol;
    a.delta = p.reliable_delta > 0 ? p.reliable_delta : 0.1;
    a.max_iter = p.max_iter;
    return a;
}

template <class T, bool MIXED>
static void enqueue_iters(const Batch& B, const BArgs& a, int k0, int k1, int t0, int nt, int iters, const double* b,
                          double* x, hipStream_t s) {
    T* r = B.r.as<T>();
    T* p = B.p.as<T>();
    T* q = B.q.as<T>();
    float* xl = B.xlo.as<float>();
    const int ns = k1 - k0;
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL((k_b_k1<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, (const T*)p, q, r, b, (const double*)x);
        hipLaunchKernelGGL((k_b_finish<2, OpF1>), dim3(ns), dim3(kBlock), 0, s, a, k0, OpF1{});
        hipLaunchKernelGGL((k_b_k2<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, (const T*)p, (const T*)q, r, x, xl);
        hipLaunchKernelGGL((k_b_finish<2, OpF2<MIXED>>), dim3(ns), dim3(kBlock), 0, s, a, k0, OpF2<MIXED>{});
        hipLaunchKernelGGL((k_b_k3<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, t0, p, (const T*)r, x, xl);
    }
    HSPS_CHECK_LAUNCH();
}

static void ensure_work(Batch& B, int prec) {
    const size_t n = (size_t)std::max(B.N, 1);
    if (!B.st.p) {
        B.st = DevBuf(B.ctx, sizeof(BState) * (size_t)std::max(B.S, 1));
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

static void launch_chunk(Batch& B, const BArgs& a, int w, int chunk, const double* b, double* x, int prec,
                         hipStream_t s, bool use_graph) {
    const int k0 = B.waves[w].first, k1 = B.waves[w].second;
    const int t0 = B.h_sys_t0[k0], nt = B.h_sys_t1[k1 - 1] - t0;
    if (nt <= 0) return;
    auto direct = [&]() {
        if (prec == HSPS_MIXED) enqueue_iters<float, true>(B, a, k0, k1, t0, nt, chunk, b, x, s);
        else enqueue_iters<double, false>(B, a, k0, k1, t0, nt, chunk, b, x, s);
    };
    if (!use_graph) { direct(); return; }
    Batch::G* g = nullptr;
    for (auto& c : B.graphs)
        if (c.wave == w && c.prec == prec && c.chunk == chunk) g = &c;
    if (!g) {
        hipGraph_t graph = nullptr;
        HSPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            direct();
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HSPS_HIP(hipStreamEndCapture(s, &graph));
        Batch::G ng;
        ng.wave = w;
        ng.prec = prec;
        ng.chunk = chunk;
        const hipError_t e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HSPS_HIP(e);
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
    // small systems: one persistent workgroup each, on the aux stream
    const int n_small = B.S - B.n_large;
    HSPS_HIP(hipEventRecord(B.fork, s));
    HSPS_HIP(hipStreamWaitEvent(B.aux, B.fork, 0));
    if (n_small > 0) {
        if (prec == HSPS_MIXED)
            hipLaunchKernelGGL((k_b_small<float, true>), dim3(n_small), dim3(kSmallThreads), 0, B.aux, a, B.n_large, b, x,
                               B.r.as<float>(), B.p.as<float>(), B.q.as<float>(), B.xlo.as<float>());
        else
            hipLaunchKernelGGL((k_b_small<double, false>), dim3(n_small), dim3(kSmallThreads), 0, B.aux, a, B.n_large, b,
                               x, B.r.as<double>(), B.p.as<double>(), B.q.as<double>(), (float*)nullptr);
        HSPS_CHECK_LAUNCH();
    }
    HSPS_HIP(hipEventRecord(B.join, B.aux));
    // large systems: tiled waves on the main stream
    if (B.T > 0) {
        hipLaunchKernelGGL(k_b_init1, dim3(B.T), dim3(kBlock), 0, s, a, 0, b, (const double*)x, B.r64.as<double>());
        hipLaunchKernelGGL((k_b_finish<2, OpInit1>), dim3(B.n_large), dim3(kBlock), 0, s, a, 0, OpInit1{});
        if (prec == HSPS_MIXED)
            hipLaunchKernelGGL((k_b_init2<float, true>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                               (const double*)B.r64.as<double>(), B.r.as<float>(), B.p.as<float>(), B.xlo.as<float>());
        else
            hipLaunchKernelGGL((k_b_init2<double, false>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                               (const double*)B.r64.as<double>(), B.r.as<double>(), B.p.as<double>(), (float*)nullptr);
        hipLaunchKernelGGL((k_b_finish<1, OpInit2>), dim3(B.n_large), dim3(kBlock), 0, s, a, 0, OpInit2{});
        HSPS_CHECK_LAUNCH();
    }
    const int chunk = std::max(1, std::min(p.check_every > 0 ? p.check_every : 32, 64));
    std::vector<BState> host(B.S);
    BState* pin = nullptr;
    HSPS_HIP(hipHostMalloc((void**)&pin, sizeof(BState) * (size_t)std::max(B.S, 1) * 2, hipHostMallocDefault));
    hipEvent_t ev[2] = {nullptr, nullptr};
    try {
        HSPS_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HSPS_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        for (int w = 0; w < (int)B.waves.size(); ++w) {
            const int k0 = B.waves[w].first, k1 = B.waves[w].second;
            const int ns = k1 - k0;
            int k = 0, chunks = 0;
            bool pending = false;
            const int cap = (p.max_iter + chunk - 1) / chunk + 4;
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
        if (B.T > 0) {
            if (prec == HSPS_MIXED) {
                // flush every large system's rows (systems are not contiguous in row order)
                for (int k = 0; k < B.n_large; ++k) {
                    const int sid = B.order[k];
                    const int a0 = B.h_row_off[sid], a1 = B.h_row_off[sid + 1];
                    hipLaunchKernelGGL(k_b_flush_range, dim3(grid_for(a1 - a0)), dim3(kBlock), 0, s, a0, a1, x,
                                       B.xlo.as<float>());
                }
            }
            hipLaunchKernelGGL(k_b_final, dim3(B.T), dim3(kBlock), 0, s, a, 0, b, (const double*)x);
            hipLaunchKernelGGL((k_b_finish<1, OpFinal>), dim3(B.n_large), dim3(kBlock), 0, s, a, 0, OpFinal{});
            HSPS_CHECK_LAUNCH();
        }
        HSPS_HIP(hipStreamWaitEvent(s, B.join, 0));
        HSPS_HIP(hipMemcpyAsync(pin, B.st.p, sizeof(BState) * B.S, hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
        std::memcpy(host.data(), pin, sizeof(BState) * B.S);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(B.aux);
        (void)hipHostFree(pin);
        if (ev[0]) (void)hipEventDestroy(ev[0]);
        if (ev[1]) (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipHostFree(pin);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return host;
}

// Project autotuner: times a few chunks of the tiled waves for each tile size
// on a scratch copy of x and keeps the fastest.
static void autotune(Batch& B, const double* b, const double* x0, const hsps_params& p) {
    if (B.n_large == 0) {
        build_tiles(B, 256);
        return;
    }
    const int cands[] = {256, 512, 1024};
    hipStream_t s = B.ctx->stream;
    DevBuf xs(B.ctx, (size_t)std::max(B.N, 1) * sizeof(double));
    hsps_params q = p;
    q.check_every = 16;
    q.max_iter = 48;  // bounds the small-system kernel during tuning
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
            HSPS_HIP(hipStreamCreateWithFlags(&B->aux, hipStreamNonBlocking));
            HSPS_HIP(hipEventCreateWithFlags(&B->join, hipEventDisableTiming));
            HSPS_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
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
            B->n_large = 0;
            for (int k = 0; k < S; ++k)
                if (row_off[B->order[k] + 1] - row_off[B->order[k]] > kSmallRows) B->n_large = k + 1;
            std::vector<int> r0(S), r1(S);
            for (int k = 0; k < S; ++k) {
                r0[k] = row_off[B->order[k]];
                r1[k] = row_off[B->order[k] + 1];
            }
            B->sys_row0 = DevBuf(ctx, sizeof(int) * (size_t)std::max(S, 1));
            B->sys_row1 = DevBuf(ctx, sizeof(int) * (size_t)std::max(S, 1));
            if (S) {
                HSPS_HIP(hipMemcpy(B->sys_row0.p, r0.data(), sizeof(int) * S, hipMemcpyHostToDevice));
                HSPS_HIP(hipMemcpy(B->sys_row1.p, r1.data(), sizeof(int) * S, hipMemcpyHostToDevice));
            }
            build_tiles(*B, 256);
            B->tile_rows = 0;  // not tuned yet
            // Wave working-set budget (MB).  Measured on MI355X (512 systems,
            // 1.9 TB of CG traffic): one wave beats MALL-sized (<=192 MB)
            // waves — the fixed per-iteration launch cost of extra waves
            // outweighs the cache hits — so the default is a single wave.
            const char* wmb = std::getenv("HSPS_BATCH_WAVE_MB");
            const size_t budget_mb = wmb ? (size_t)std::max(8, std::atoi(wmb)) : ((size_t)1 << 20);
            build_waves(*B, budget_mb << 20);
        } catch (...) {
            delete B;
            throw;
        }
        ctx_retain(ctx);
        *out = B;
        return HSPS_OK;
    });
}

int hsps_batch_destroy(hsps_batch B) {
    return guarded([&] {
        if (B) {
            Ctx* c = B->ctx;
            (void)hipSetDevice(c->device);
            (void)hipStreamSynchronize(c->stream);
            (void)hipStreamSynchronize(B->aux);
            delete B;
            ctx_release(c);
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
            if (p.batch_tile_rows != B->tile_rows) build_tiles(*B, std::min(p.batch_tile_rows, 4096));
        } else if (B->tile_rows == 0) {
            ensure_work(*B, p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64);
            autotune(*B, b, x, p);
        }
        // HSPS_BATCH_NO_GRAPH=1 launches the iteration chunks directly (for
        // tracing tools that do not follow graph launches).
        static const bool no_graph = std::getenv("HSPS_BATCH_NO_GRAPH") != nullptr;
        std::vector<BState> st = run_batch(*B, b, x, p, !no_graph);
        int n_conv = 0;
        for (int k = 0; k < B->S; ++k) {
            const int s = B->order[k];
            const BState& S = st[k];
            if (iters) iters[s] = S.iter;
            if (conv) conv[s] = S.conv;
            if (true_res) true_res[s] = S.true_rel;
            n_conv += S.conv;
        }
        return n_conv == B->S ? 1 : 0;
    });
}

}  // extern "C"
