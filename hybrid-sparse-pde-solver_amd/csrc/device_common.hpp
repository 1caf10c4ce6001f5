// =============================================================================
// device_common.hpp — gfx950 device building blocks: wave64 / workgroup /
// two-phase grid reductions (partials + finish kernel), and the CSR
// row-block SpMV template with pluggable epilogues.
// =============================================================================
#pragma once

#include <hip/hip_runtime.h>

#include "hsps_internal.hpp"

namespace hsps_dev {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum NV values over a workgroup of BS threads; the result is valid in
// thread 0.  `lds` must hold NV * (BS/64) doubles.  Ends with a barrier-free result
// in thread 0; callers that reuse `lds` must barrier first.
template <int NV, int BS = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
    constexpr int W = BS / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k * W + wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < W; ++w) s += lds[k * W + w];
            v[k] = s;
        }
    }
}

// Two-phase grid reductions.  Phase 1 (in the producing kernel): every
// workgroup writes its NV partial sums to partials[k * gridDim.x + block].
// Phase 2 (a one-workgroup finish kernel launched next on the same stream):
// sums the partials in a fixed order — deterministic for a given grid — and
// runs the caller's scalar logic.  No cross-workgroup signalling inside a
// kernel: on gfx950 an agent-scope release/acquire per workgroup costs an L2
// writeback + invalidate (buffer_wbl2 / buffer_inv), which destroys the L2
// reuse of the SpMV gathers; the kernel boundary gives visibility for free.
template <int NV>
__device__ __forceinline__ void block_partial(double (&v)[NV], double* partials) {
    __shared__ double red[NV * (kBlock / 64)];
    block_sum<NV>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
}

// Phase 2 helper: sum nparts partials (NV planes) with the whole workgroup;
// totals valid in thread 0.
template <int NV>
__device__ __forceinline__ void finish_sum(const double* partials, int nparts, double (&tot)[NV]) {
    __shared__ double red[NV * (kBlock / 64)];
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += partials[(size_t)k * nparts + b];
    }
    block_sum<NV>(acc, red);
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = acc[k];
}

// One-workgroup finish kernel: op.skip() gates it (uniform), op(tot) runs in
// thread 0.
template <int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_finish(const double* partials, int nparts, Op op) {
    if (op.skip()) return;
    double tot[NV];
    finish_sum<NV>(partials, nparts, tot);
    if (threadIdx.x == 0) op(tot);
}

template <int NV, class Op>
inline void launch_finish(const double* partials, int nparts, const Op& op, hipStream_t s) {
    hipLaunchKernelGGL((k_finish<NV, Op>), dim3(1), dim3(kBlock), 0, s, partials, nparts, op);
}

// Finish op that stores the NV totals (optionally behind a skip flag).
template <int NV>
struct OpStore {
    double* out;
    const int* flag = nullptr;
    __device__ bool skip() const { return flag && *flag; }
    __device__ void operator()(const double* tot) const {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = tot[k];
    }
};

template <class T> struct AccOf;
template <> struct AccOf<double> { using type = double; };
template <> struct AccOf<float> { using type = float; };

// CSR row-block SpMV.  Workgroup b owns rows [blocks[b], blocks[b+1]).
//  * staged path (block nnz <= kLdsNnz for FP64 sums, kLdsNnz32 for FP32,
//    on the matching block table): the workgroup streams val/col
//    contiguously, writes val*x[col] products to LDS, then sums each row from
//    LDS with tpr lanes per row (tpr = pow2 <= 256/rows, <= 64).
//  * long-row path (one row with more nonzeros than that): the whole workgroup strides
//    the row and reduces.
// Epi supplies: skip(), row(row, sum, red[]), NV, and the partials buffer
// (NV * n_blocks doubles) when NV > 0; a finish kernel consumes the partials.
template <typename V, typename T, class Epi>
__global__ void __launch_bounds__(kBlock)
k_spmv(const int* __restrict__ blocks, const int* __restrict__ rp, const int* __restrict__ ci,
       const V* __restrict__ val, const T* __restrict__ x, Epi epi) {
    using Acc = typename AccOf<T>::type;
    constexpr int kCap = sizeof(Acc) == 8 ? kLdsNnz : kLdsNnz32;  // 16 KB of products either way
    if (epi.skip()) return;
    __shared__ Acc prod[kCap];
    const int tid = threadIdx.x;
    const int r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
    const int k0 = rp[r0], k1 = rp[r1];
    double red[Epi::NV > 0 ? Epi::NV : 1];
#pragma unroll
    for (int k = 0; k < (Epi::NV > 0 ? Epi::NV : 1); ++k) red[k] = 0.0;
    if (k1 - k0 > kCap) {
        Acc s = 0;
        for (int k = k0 + tid; k < k1; k += kBlock) s += (Acc)val[k] * (Acc)x[ci[k]];
        double v1[1] = {(double)s};
        __shared__ double lred[kBlock / 64];
        block_sum<1>(v1, lred);
        if (tid == 0) epi.row(r0, (Acc)v1[0], red);
    } else {
        // kCap / kBlock products per thread: every column and value load is
        // issued, then every x gather, then the LDS stores, so a wave keeps
        // all its loads in flight at once instead of two per product.
        // The trip count nu is uniform over the workgroup (a scalar branch,
        // no exec-mask split); lanes past the block's last nonzero re-read
        // that nonzero and park a product in an LDS slot no row reads.
        constexpr int U = kCap / kBlock;
        const int nu = (k1 - k0 + kBlock - 1) / kBlock;
        unsigned col[U];
        V vv[U];
        T xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            col[u] = 0;
            vv[u] = V(0);
            xv[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < nu) {
                const int k = min(k0 + tid + u * kBlock, k1 - 1);
                col[u] = (unsigned)ci[k];
                vv[u] = val[k];
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < nu) xv[u] = x[col[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < nu) prod[tid + u * kBlock] = (Acc)vv[u] * (Acc)xv[u];
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
            if (l == 0) epi.row(row, s, red);
        }
    }
    if constexpr (Epi::NV > 0) block_partial<Epi::NV>(red, epi.partials);
}

// ---- basic epilogues ------------------------------------------------------------
struct SkipFlag {
    const int* flag = nullptr;
    __device__ bool skip() const { return flag && *flag; }
};

template <typename T>
struct EpiStore : SkipFlag {
    static constexpr int NV = 0;
    T* y;
    double* partials = nullptr;
    template <class A> __device__ void row(int i, A s, double*) const { y[i] = (T)s; }
};

template <typename T>
struct EpiAdd : SkipFlag {
    static constexpr int NV = 0;
    T* y;
    double* partials = nullptr;
    template <class A> __device__ void row(int i, A s, double*) const { y[i] = (T)((A)y[i] + s); }
};

// r = b - A x; with NORM, ||r||^2 partials for a finish kernel
template <typename T, bool NORM>
struct EpiResidual : SkipFlag {
    static constexpr int NV = NORM ? 1 : 0;
    const T* b;
    T* r;
    double* partials = nullptr;
    template <class A> __device__ void row(int i, A s, double* red) const {
        const T v = (T)((A)b[i] - s);
        r[i] = v;
        if constexpr (NORM) red[0] += (double)v * (double)v;
    }
};

// xo = xi + omega * dinv * (rhs - A xi)   (amg.cpp:242-254, one sweep)
template <typename T>
struct EpiSmooth : SkipFlag {
    static constexpr int NV = 0;
    const T* dinv;
    const T* rhs;
    const T* xi;
    T* xo;
    T omega;
    double* partials = nullptr;
    template <class A> __device__ void row(int i, A s, double*) const {
        xo[i] = (T)((A)xi[i] + (A)omega * (A)dinv[i] * ((A)rhs[i] - s));
    }
};

// Row blocks the SpMV over vectors of type T runs on (the count of partial
// sums an epilogue with NV > 0 leaves for its finish kernel).
template <typename T> inline int spmv_blocks(const Csr& A) {
    return sizeof(typename AccOf<T>::type) == 8 ? A.n_blocks : A.n_blocks32;
}

template <typename V, typename T, class Epi>
inline void launch_spmv(const Csr& A, const V* val, const T* x, const Epi& epi, hipStream_t s) {
    const int nb = spmv_blocks<T>(A);
    if (nb == 0) return;
    const int* blk = sizeof(typename AccOf<T>::type) == 8 ? A.blocks.as<int>() : A.blocks32.as<int>();
    hipLaunchKernelGGL((k_spmv<V, T, Epi>), dim3(nb), dim3(kBlock), 0, s, blk, A.rp.as<int>(), A.ci.as<int>(), val,
                       x, epi);
    HSPS_CHECK_LAUNCH();
}

// Grid size for element-wise / reduction kernels over n entries.
inline int grid_for(int n, int per_thread = 4) {
    long g = ((long)n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (int)g;
}

}  // namespace hsps_dev
