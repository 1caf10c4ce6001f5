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
//    (default: one wave).  Per iteration, over the wave's tiles:
//      K1  p' = D^-1 r + beta p (two p buffers), q = A p', p'.q partials
//          (FLUSH state: FP64 true residual r = b - A (x64 + x_lo))
//      F1  per-system finish: alpha (or breakdown / reliable-update decisions)
//      K2  x += alpha p', r -= alpha q, r.r / r.z partials
//          (after a FLUSH: x64 += x_lo)
//      F2  per-system finish: rel, rho, beta, convergence
//    F1 and F2 run inside K1 and K2: the system's workgroup that starts last
//    sums the tiles' partials and decides, so an iteration is two launches.
//    The chunk of iterations is captured once per wave in a hipGraph and
//    replayed; the host polls the wave's states one chunk behind.
// Systems on <= 8 diagonals are stored by diagonal (constant-coefficient
// stencils as one value per diagonal + a presence byte per row), so their
// iterations move ~34 B per row (FP32) instead of ~96 B with CSR.
// The two classes run concurrently on two streams.
// Mixed precision: FP32 matrix values and Krylov vectors, FP64 reductions and
// scalars, x accumulated in FP64 through an FP32 low-order part flushed at
// every reliable update (|r| < delta * max|r| since the last one, or |r| <
// tol to confirm convergence against the FP64 true residual).
// =============================================================================
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <type_traits>

#include "device_common.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr int kSmallRows = 20000;   // persistent one-workgroup path up to this size
constexpr int kSmallThreads = 1024;
constexpr int kMaxDiags = 8;        // DIA storage when a system's nonzeros lie on <= 8 diagonals

// LDS staging window of K1 / K2 (dynamic LDS, in doubles): 16 KB unless a
// system's halo needs more (measured on MI355X with halos up to 512 rows, a
// 32 KB window costs more in occupancy than it saves in halo re-reads); at
// most 64 KB.  A chunk fills the 16 KB window, or takes 2 halo rows when the
// halo is wider (as far as 64 KB allows).
constexpr int kStageMinDoubles = 2048, kStageMaxDoubles = 8192;

// Rows per LDS chunk of a diagonal-stored system with halo h (largest |offset|)
// and element size es: a function of the system alone, so that its rows are
// reduced in the same order whatever window a launch provides.
__host__ __device__ constexpr int stage_halo(int h, int es) { return (h + 16 / es - 1) / (16 / es) * (16 / es); }
__host__ __device__ constexpr int stage_chunk(int h, int es) {
    const int G = 16 / es, hr = stage_halo(h, es);
    const int fill = kStageMinDoubles * 8 / es - 2 * hr - 2 * G;  // what the 16 KB window leaves
    const int room = kStageMaxDoubles * 8 / es - 2 * hr - 2 * G;  // what the 64 KB window leaves
    const int want = 2 * hr < room ? 2 * hr : room;
    const int c = (fill > want ? fill : want) / G * G;
    return c > 0 ? c : 0;
}
// LDS elements the chunk and its halo window take (more than any window when
// no chunk fits).
__host__ __device__ constexpr int stage_need(int h, int es) {
    return stage_chunk(h, es) >= 16 / es ? 2 * stage_halo(h, es) + 2 * (16 / es) + stage_chunk(h, es) : 1 << 30;
}


// NORMAL: K1 updates p and forms q = A p, K2 updates x and r.
// FLUSH: K1 replaces r by the FP64 true residual b - A (x64 + x_lo).
// FLUSHED: after a FLUSH, K2 folds x_lo into x64; K1 then proceeds as NORMAL.
enum : int { BS_NORMAL = 0, BS_FLUSH = 1, BS_FLUSHED = 2 };
enum : int { BR_NONE = 0, BR_RELIABLE = 1, BR_CHECK = 2, BR_RESTART = 3, BR_STALLED = 4 };
// storage of a system inside the batch
enum : int { SK_CSR = 0, SK_DIA = 1, SK_STENCIL = 2 };

struct BState {
    // the fields every K1 / K2 workgroup reads come first: one 32-byte load (BHot)
    double alpha, beta;
    int halt, flag;
    int pcur;  // which of the two p buffers holds the current search direction
    int seq;   // finishes so far: K1 / K2 launches tag their partials seq + 1
    double rho, rho_old, bnorm, rel, rmax, true_rel;
    int done, reason, reliable, restarts, conv;
    int restarts_pq;  // restarts after a p.q breakdown (the rest follow a rho breakdown)
    double r0;  // initial ||r|| (absolute)
    int iter;
};
static_assert(sizeof(BState) == 120, "BState layout");
struct BHot {
    double alpha, beta;
    int halt, flag, pcur, seq;
};
static_assert(offsetof(BState, alpha) == offsetof(BHot, alpha) && offsetof(BState, seq) == offsetof(BHot, seq),
              "BHot is the head of BState");
// The head of a system's state in one load (the compiler would otherwise
// issue each field's load after the branch on the one before).
__device__ __forceinline__ BHot load_hot(const BState* S) {
    BHot h;
    __builtin_memcpy(&h, S, sizeof(BHot));
    return h;
}

// Per-tile record of the 16-byte five-point path (a stencil-stored system with
// at most five diagonals, a scalar Jacobi diagonal and a staging window that
// fits the launch's LDS): everything K1 / K2 read before their first vector
// load, in one 128-byte record, so that a workgroup waits for two dependent
// scalar loads (this record, then its system's state) instead of one per
// table.  vec = 1 / 2: the five-point kernels with kVecNarrow / kVecWide
// threads; vec = 0 sends the tile down the general paths, which read the
// tables themselves.  The iteration kernels read the records of the tiles
// whose systems are still iterating from a compacted copy (BArgs::tva*),
// which the host rebuilds as systems halt.  vec = 3: a CSR system none of
// whose tiles exceeds the LDS product stage, run by k_b_k1c / k_b_k2c (every
// tile of a system takes the same kernels, so its fused finish finds all of
// them in one launch).
template <class T>
struct alignas(128) TileVec {
    int r0, r1, sys, vec;  // rows [r0, r1) of internal system sys
    int row0, last, H, C;  // the system's rows [row0, last]; staging halo and chunk (stage_halo / stage_chunk)
    int off[5];            // diagonal offsets, ascending (padding: offset 0 with value 0, never present)
    int tile;              // the tile's index (its partials slot)
    int fin;               // in an active list: the system's workgroup the dispatcher starts last (runs the finish)
    int cdok;              // D^-1 is the scalar cd (else the dinv vector)
    T uv[5];               // the diagonals' values
    T cd;                  // D^-1
};

struct BArgs {
    const TileVec<float>* tv32;
    const TileVec<double>* tv64;
    const TileVec<float>* tva32;   // records of the active tiles (one per workgroup of K1 / K2)
    const TileVec<double>* tva64;
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
    // Diagonal storage of systems whose nonzeros lie on <= kMaxDiags diagonals
    // (column = row + off[s*kMaxDiags+d], offsets ascending):
    //  SK_DIA:     value of diagonal d at row li of system s in
    //              dia[base[s] + d * n_s + li] (zero where absent);
    //  SK_STENCIL: every diagonal carries one value uv[s*kMaxDiags+d]; bit d of
    //              umask[row] says whether the row holds that diagonal.
    const int* dia_kind;       // SK_* per system
    const int* dia_nd;         // diagonals per system
    const int* dia_off;        // S * kMaxDiags offsets
    const long long* dia_base;
    const double* dia64;
    const float* dia32;
    const double* uv64;        // S * kMaxDiags
    const float* uv32;
    const unsigned char* umask;  // N
    // Jacobi with one diagonal value for the whole system: D^-1 = cd[s]
    const int* cd_ok;
    const double* cd64;
    const float* cd32;
    int stage_doubles;  // dynamic LDS of K1 / K2, in doubles (>= kLdsNnz for the CSR path)
    BState* st;
    double* partials;
    unsigned long long* tagp;  // [2 phases][T tiles][2 values][2 words] tagged partials of K1 / K2
    int T;  // total tiles
    double tol, delta;
    int max_iter;
};

template <class T> struct Prec;
template <> struct Prec<double> {
    static __device__ const TileVec<double>* tv(const BArgs& B) { return B.tv64; }
    static __device__ const TileVec<double>* tva(const BArgs& B) { return B.tva64; }
    static __device__ const double* val(const BArgs& B) { return B.v64; }
    static __device__ const double* dinv(const BArgs& B) { return B.dinv64; }
    static __device__ const double* dia(const BArgs& B) { return B.dia64; }
    static __device__ const double* uv(const BArgs& B) { return B.uv64; }
    static __device__ const double* cd(const BArgs& B) { return B.cd64; }
};
template <> struct Prec<float> {
    static __device__ const TileVec<float>* tv(const BArgs& B) { return B.tv32; }
    static __device__ const TileVec<float>* tva(const BArgs& B) { return B.tva32; }
    static __device__ const float* val(const BArgs& B) { return B.v32; }
    static __device__ const float* dinv(const BArgs& B) { return B.dinv32; }
    static __device__ const float* dia(const BArgs& B) { return B.dia32; }
    static __device__ const float* uv(const BArgs& B) { return B.uv32; }
    static __device__ const float* cd(const BArgs& B) { return B.cd32; }
};

// Row access of a diagonal-stored system, loaded once per workgroup.  Rows
// are read without column indices; every load is coalesced across rows.
template <class T>
struct Stencil {
    int kind, nd, row0, n_s;
    int halo;  // max |offset|
    int off[kMaxDiags];  // constant-indexed copies (unrolled paths)
    T uv[kMaxDiags];
    const int* __restrict__ offg;  // the same in memory (loop over diagonals)
    const T* __restrict__ uvg;
    const T* __restrict__ dia;
    const unsigned char* __restrict__ mask;
    __device__ __forceinline__ Stencil(const BArgs& B, int s) {
        kind = B.dia_kind[s];
        nd = B.dia_nd[s];
        row0 = B.sys_row0[s];
        n_s = B.sys_row1[s] - row0;
        dia = Prec<T>::dia(B) + B.dia_base[s];
        mask = B.umask;
        const T* u = Prec<T>::uv(B) + (size_t)s * kMaxDiags;
        offg = B.dia_off + s * kMaxDiags;
        uvg = u;
        halo = 0;
#pragma unroll
        for (int d = 0; d < kMaxDiags; ++d) {
            off[d] = B.dia_off[s * kMaxDiags + d];
            uv[d] = u[d];
            if (d < nd) halo = max(halo, abs(off[d]));
        }
    }
    // f(integral_constant<ND>): ND = 5 unrolls the row (1D and 2D five-point
    // stencils; padding diagonals contribute nothing); ND = 0 walks the
    // diagonals in a loop (wider stencils), keeping the kernels' register
    // budget at what the five-point case needs.
    template <class F>
    __device__ __forceinline__ void dispatch(F&& f) const {
        if (nd <= 5) f(std::integral_constant<int, 5>{});
        else f(std::integral_constant<int, 0>{});
    }
    // Row i of A x, x(j) = xv(j): sum over the row's diagonals in ascending
    // column order of A(i, j) * x(j) with rounded products (no fma
    // contraction), i.e. exactly the CSR loop sum += v * x in column order.
    // Every x(j) is read unconditionally at an index clamped into the system
    // and masked afterwards, so the loads of a row are all in flight at once.
    template <int ND, class XF>
    __device__ __forceinline__ T row(int i, XF&& xv) const {
#pragma clang fp contract(off)
        const int li = i - row0, last = row0 + n_s - 1;
        if constexpr (ND == 0) {
            const unsigned m = kind == SK_STENCIL ? mask[i] : 0u;
            T sum = 0;
            for (int d = 0; d < nd; ++d) {
                const int od = offg[d];
                const int lj = li + od;
                if (kind == SK_STENCIL ? ((m >> d) & 1u) != 0 : (lj >= 0 && lj < n_s)) {
                    const T a = kind == SK_STENCIL ? uvg[d] : dia[(size_t)d * n_s + li];
                    const T t = a * (T)xv(i + od);
                    sum = sum + t;
                }
            }
            return sum;
        } else {
            T xs[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) xs[d] = (T)xv(min(max(i + off[d], row0), last));
            T sum = 0;
            if (kind == SK_STENCIL) {
                const unsigned m = mask[i];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const T t = uv[d] * xs[d];
                    sum = ((m >> d) & 1u) ? sum + t : sum;
                }
            } else {
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const int lj = li + off[d];
                    const bool ok = d < nd && lj >= 0 && lj < n_s;
                    const T t = dia[(size_t)(ok ? d : 0) * n_s + li] * xs[d];
                    sum = ok ? sum + t : sum;
                }
            }
            return sum;
        }
    }
};

// Rows r0..r1 of A x for a diagonal-stored system, with x staged in LDS:
// chunks of C = cap - 2 halo rows; for each chunk x(j) = xv(j) is evaluated
// once per row of the chunk plus its halo (coalesced, four rows in flight per
// thread) into lds, own(j, x_j) is called for the chunk's own rows, then
// f(i, x_i, (A x)_i) once per row with the row read from LDS.  Callers check
// that stage_need(halo) <= cap.
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;

template <int ND, class T, class XF, class OWN, class F>
__device__ __forceinline__ void stencil_staged(const Stencil<T>& A, int r0, int r1, lds_ptr<T> lds, int cap,
                                               XF&& xv, OWN&& own, F&& f) {
    constexpr int U = 4;
    const int tid = threadIdx.x;
    const int C = stage_chunk(A.halo, (int)sizeof(T));  // 2 halo + C <= stage_need <= cap (checked by the caller)
    const int lo = A.row0, hi = A.row0 + A.n_s;
    for (int c0 = r0; c0 < r1; c0 += C) {
        const int c1 = min(r1, c0 + C);
        const int base = c0 - A.halo;
        const int s0 = max(base, lo), s1 = min(c1 + A.halo, hi);
        __syncthreads();  // the previous chunk's readers are done with lds
        for (int j0 = s0 + tid; j0 < s1; j0 += U * kBlock) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = xv(min(j0 + u * kBlock, s1 - 1));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + u * kBlock;
                if (j < s1) {
                    lds[j - base] = v[u];
                    if (j >= c0 && j < c1) own(j, v[u]);
                }
            }
        }
        __syncthreads();
        for (int i = c0 + tid; i < c1; i += kBlock)
            f(i, lds[i - base], A.template row<ND>(i, [&](int j) { return lds[j - base]; }));
    }
}

// ---- 16-byte groups: G = 16 / sizeof(T) consecutive rows per lane ---------------------------
// The iteration kernels of five-point stencil systems move every vector in
// 16-byte loads and stores (four FP32 or two FP64 rows per lane), stage the
// x values a row's stencil needs in LDS once per chunk, and read the presence
// bytes of a group as one word.  Work vectors and the presence bytes are
// allocated with kPad rows of slack, so a group that runs past the last row
// stays inside the allocation; rows outside [c0, c1) are neither written nor
// reduced.
template <class T>
struct alignas(16) Pack {
    static constexpr int G = 16 / (int)sizeof(T);
    T e[G];
};
constexpr int kPad = 16;

template <class T>
__device__ __forceinline__ Pack<T> ld16(const T* p) { return *reinterpret_cast<const Pack<T>*>(p); }
template <class T>
__device__ __forceinline__ void st16(T* p, const Pack<T>& v) { *reinterpret_cast<Pack<T>*>(p) = v; }
// Streaming 16-byte store (nontemporal hint) for K1's new direction: it is
// read back by the next two kernels only, and keeping it out of the caches'
// way leaves more of the 256 MB Infinity Cache to r and x_lo.
template <class T>
struct Vec16 {
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};
template <class T>
__device__ __forceinline__ void st16_stream(T* p, const Pack<T>& v) {
    typename Vec16<T>::type w;
#pragma unroll
    for (int e = 0; e < Pack<T>::G; ++e) w[e] = v.e[e];
    __builtin_nontemporal_store(w, reinterpret_cast<typename Vec16<T>::type*>(p));
}
// 16-byte LDS store / load of a group (its offset is a multiple of the
// group): one ds_write_b128 / ds_read_b128 instead of word pairs that lie
// four words apart across lanes, on the same banks
template <class T>
__device__ __forceinline__ void st_lds16(__attribute__((address_space(3))) T* p, const Pack<T>& v) {
    typename Vec16<T>::type w;
#pragma unroll
    for (int e = 0; e < Pack<T>::G; ++e) w[e] = v.e[e];
    *reinterpret_cast<__attribute__((address_space(3))) typename Vec16<T>::type*>(p) = w;
}
template <class T>
__device__ __forceinline__ Pack<T> ld_lds16(const __attribute__((address_space(3))) T* p) {
    const typename Vec16<T>::type w = *reinterpret_cast<const __attribute__((address_space(3))) typename Vec16<T>::type*>(p);
    Pack<T> v;
#pragma unroll
    for (int e = 0; e < Pack<T>::G; ++e) v.e[e] = w[e];
    return v;
}
// staging: groups per lane whose loads are issued together (K2: a 2048-row
// tile with 512-row halos in one round; K1 holds its first presence bytes
// across the staging, and two pairs of groups keep it within 64 registers)
constexpr int kStageU = 3, kStageK1 = 2;
template <class T>
__device__ __forceinline__ unsigned mask_word(const unsigned char* m, int g) {
    if constexpr (Pack<T>::G == 4) return *reinterpret_cast<const unsigned*>(m + g);
    else return *reinterpret_cast<const unsigned short*>(m + g);
}

// Row i of a five-point stencil system from x staged in LDS (lds[j - base]
// holds x_j), presence bits mb: the same sum, in the same order and rounding,
// as Stencil::row<5>.  Every index stays inside the staging window (|off| <=
// halo <= H and the rows lie in [a0, c1 + G - 1)); a neighbour outside the
// system (whose LDS slot may hold anything) has its presence bit clear and is
// selected away, so no index needs clamping and a group's rows read
// consecutive LDS words.
template <class T>
__device__ __forceinline__ T row_lds(const TileVec<T>& D, int i, unsigned mb, lds_ptr<T> lds, int base) {
#pragma clang fp contract(off)
    T sum = 0;
#pragma unroll
    for (int d = 0; d < 5; ++d) {
        const T t = D.uv[d] * lds[i + D.off[d] - base];
        sum = ((mb >> d) & 1u) ? sum + t : sum;
    }
    return sum;
}

// Chunk geometry of the vectorised stencil paths: chunks of C rows (a
// multiple of G) whose staged window [g0, g1) (groups aligned to G, rows
// inside the system) holds every column the chunk's rows touch.
template <class T>
struct VecChunk {
    static constexpr int G = Pack<T>::G;
    int c0, c1, a0, base, g0, g1;
    __device__ __forceinline__ VecChunk(const TileVec<T>& D, int c) {
        c0 = c;
        c1 = min(D.r1, c + D.C);
        a0 = c0 & ~(G - 1);
        base = a0 - D.H;
        g0 = max(base, D.row0 & ~(G - 1));
        g1 = min(c1 + D.H, D.last + 1);
    }
};
constexpr int kRowsU = 8;  // K1's stencil phase: rows per lane whose presence bytes load together
// Workgroup sizes of the five-point kernels: 256 threads, or 512 for a
// system whose staging window exceeds 16 KB (a wide halo: at 32-64 KB of LDS a
// CU holds 2-4 workgroups, and 512-thread ones keep twice the waves in flight;
// Poisson 2048^2, one system, mixed PCG: 82 -> 76 ms; the 16 KB windows of the
// 512-system batch run faster at 256).  The size is a property of the system
// (TileVec::vec = 1 or 2), so a system gives the same bits alone and in any batch.
constexpr int kVecNarrow = kBlock, kVecWide = 512;

// K1 rows [r0, r1) of a five-point stencil system: p_new = dv r + beta p_old
// by 16-byte groups into LDS, then q = A p_new from LDS and pq += p_new . q
// (one row per lane at a time), then the chunk's own rows of p_new out to pn.
// q is not stored: K2 recomputes it.
template <class T, int BS>
__device__ __forceinline__ void k1_stencil_vec(const TileVec<T>& D, const unsigned char* __restrict__ mask,
                                               lds_ptr<T> lds, const T* __restrict__ r, const T* __restrict__ po,
                                               T* __restrict__ pn, double beta, double& pq) {
    constexpr int G = Pack<T>::G;
    const int tid = threadIdx.x;
    const T dv = D.cd;
    for (int c = D.r0; c < D.r1; c += D.C) {
        const VecChunk<T> W(D, c);
        // the presence bytes of the stencil phase's first kRowsU rows per
        // lane, loaded ahead of the staging (they arrive with it instead of
        // after the barrier; 117.5 vs 118.5 ms for the 512-system solve)
        unsigned mb0[kRowsU];
#pragma unroll
        for (int u = 0; u < kRowsU; ++u) mb0[u] = W.c0 + tid + u * BS < W.c1 ? mask[W.c0 + tid + u * BS] : 0u;
        __syncthreads();  // the previous chunk's readers are done with lds
        // kStageK1 groups per lane: every load is issued before the first
        // store (vmcnt counts loads and stores in one in-order queue, so a
        // load wait placed after a store also waits for that store)
        for (int gb = W.g0 + tid * G; gb < W.g1; gb += kStageK1 * BS * G) {
            Pack<T> rv[kStageK1], pv[kStageK1];
#pragma unroll
            for (int u = 0; u < kStageK1; ++u) {
                const int g = gb + u * BS * G;
                if (g < W.g1) {
                    rv[u] = ld16(r + g);
                    pv[u] = ld16(po + g);
                }
            }
#pragma unroll
            for (int u = 0; u < kStageK1; ++u) {
                const int g = gb + u * BS * G;
                if (g >= W.g1) break;
                Pack<T> v;
#pragma unroll
                for (int e = 0; e < G; ++e) v.e[e] = (T)((double)(T)(dv * rv[u].e[e]) + beta * (double)pv[u].e[e]);
                st_lds16(lds + (g - W.base), v);
            }
        }
        __syncthreads();
        // q = A p_new and p_new . q over the chunk's rows, consecutive rows on
        // consecutive lanes (lane t takes rows c0 + t + j BS): a wave's LDS reads
        // of one stencil offset are consecutive words, free of bank conflicts
        // (four rows per lane read words four apart: 4-way conflicts, 2% slower
        // at full load).  The presence bytes of kRowsU rows are loaded together;
        // a row past the chunk is computed on the chunk's last row and dropped.
        for (int i0 = W.c0 + tid; i0 < W.c1; i0 += kRowsU * BS) {
            unsigned mb[kRowsU];
#pragma unroll
            for (int u = 0; u < kRowsU; ++u)
                mb[u] = i0 == W.c0 + tid ? mb0[u] : i0 + u * BS < W.c1 ? mask[i0 + u * BS] : 0u;
#pragma unroll
            for (int u = 0; u < kRowsU; ++u) {
                const int i = min(i0 + u * BS, W.c1 - 1);
                const T sum = row_lds(D, i, mb[u], lds, W.base);
                const double t = fma((double)lds[i - W.base], (double)sum, pq);
                pq = i0 + u * BS < W.c1 ? t : pq;
            }
        }
        // the chunk's own rows of p_new, stored only now: a store issued
        // before the stencil phase would hold up its first presence-word use
        // (vmcnt counts loads and stores in one in-order queue)
        for (int g = W.a0 + tid * G; g < W.c1; g += BS * G) {
            const Pack<T> v = ld_lds16(lds + (g - W.base));
            if (g >= W.c0 && g + G <= W.c1) {
                st16_stream(pn + g, v);
            } else {
#pragma unroll
                for (int e = 0; e < G; ++e)
                    if (g + e >= W.c0 && g + e < W.c1) pn[g + e] = v.e[e];
            }
        }
    }
}

// K2 rows [r0, r1) of a five-point stencil system: q = A p recomputed from p
// staged in LDS, then x += alpha p (x_lo in mixed mode), r -= alpha q, with
// r.r and r.z (z = dv r) partials; own-row vectors by 16-byte groups.
template <class T, bool MIXED, int BS>
__device__ __forceinline__ void k2_stencil_vec(const TileVec<T>& D, const unsigned char* __restrict__ mask,
                                               lds_ptr<T> lds, const T* __restrict__ p, T* __restrict__ r,
                                               double* __restrict__ x64, float* __restrict__ xlo, double alpha,
                                               double (&red)[2]) {
    constexpr int G = Pack<T>::G;
    using X = std::conditional_t<MIXED, float, double>;  // the x part K2 updates
    static_assert(Pack<X>::G == G, "x and the Krylov vectors group alike");
    const int tid = threadIdx.x;
    const T dv = D.cd;
    X* __restrict__ xs = MIXED ? (X*)xlo : (X*)x64;
    for (int c = D.r0; c < D.r1; c += D.C) {
        const VecChunk<T> W(D, c);
        __syncthreads();
        // staging loads kStageU groups at a time, all in flight before the first LDS store
        for (int g0 = W.g0 + tid * G; g0 < W.g1; g0 += kStageU * BS * G) {
            Pack<T> v[kStageU];
#pragma unroll
            for (int u = 0; u < kStageU; ++u)
                if (g0 + u * BS * G < W.g1) v[u] = ld16(p + g0 + u * BS * G);
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                const int g = g0 + u * BS * G;
                if (g >= W.g1) break;
                st_lds16(lds + (g - W.base), v[u]);
            }
        }
        __syncthreads();
        for (int g = W.a0 + tid * G; g < W.c1; g += BS * G) {
            const bool full = g >= W.c0 && g + G <= W.c1;
            const unsigned mw = mask_word<T>(mask, g);
            const Pack<T> rv = ld16(r + g);
            Pack<X> xv;
            if (MIXED || full) {  // x_lo is padded; the caller's x64 only read in full groups
                xv = ld16(xs + g);
            } else {
#pragma unroll
                for (int e = 0; e < G; ++e) xv.e[e] = (g + e >= W.c0 && g + e < W.c1) ? xs[g + e] : X(0);
            }
            Pack<T> ro;
            Pack<X> xo;
#pragma unroll
            for (int e = 0; e < G; ++e) {
                const int i = g + e;
                const T pi = lds[i - W.base];
                const T qi = row_lds(D, i, (mw >> (8 * e)) & 0xffu, lds, W.base);
                xo.e[e] = (X)((double)xv.e[e] + alpha * (double)pi);
                const T ri = (T)((double)rv.e[e] - alpha * (double)qi);
                ro.e[e] = ri;
                if (i >= W.c0 && i < W.c1) {
                    const double rd = (double)ri;
                    red[0] += rd * rd;
                    red[1] += rd * (double)(T)(dv * ri);
                }
            }
            if (full) {
                st16(r + g, ro);
                st16(xs + g, xo);
            } else {
#pragma unroll
                for (int e = 0; e < G; ++e)
                    if (g + e >= W.c0 && g + e < W.c1) {
                        r[g + e] = ro.e[e];
                        xs[g + e] = xo.e[e];
                    }
            }
        }
    }
}

// Row sums of a tile (rows r0..r1) of A*x with x(j) = xv(j), LDS-staged;
// f(row, sum) once per row from one lane.  Same staging rule as the
// single-system SpMV.
template <class Acc, class V, class XF, class P, class F>
__device__ __forceinline__ void tile_spmv(int r0, int r1, const int* __restrict__ rp, const int* __restrict__ ci,
                                          const V* __restrict__ val, XF&& xv, P prod, F&& f) {
    const int tid = threadIdx.x;
    const int k0 = rp[r0], k1 = rp[r1];
    if (k1 - k0 > kLdsNnz) {  // one long row
        Acc s = 0;
        for (int k = k0 + tid; k < k1; k += kBlock) s += (Acc)val[k] * (Acc)xv(ci[k]);
        double v1[1] = {(double)s};
        __shared__ double lred[kBlock / 64];
        block_sum<1>(v1, lred);
        if (tid == 0) f(r0, (Acc)v1[0]);
        return;
    }
    for (int k = k0 + tid; k < k1; k += kBlock) prod[k - k0] = (Acc)val[k] * (Acc)xv(ci[k]);
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

// FP64 rows r0..r1 of A*x for system s: diagonal-stored systems row by row
// (any tile size), CSR systems through the LDS-staged tile_spmv (tiles within
// kLdsNnz nonzeros).  f(row, sum) once per row.
template <class XF, class P, class F>
__device__ __forceinline__ void tile_rows64(const BArgs& B, int s, int r0, int r1, XF&& xv, P prod, F&& f) {
    if (B.dia_kind[s] != SK_CSR) {
        const Stencil<double> A(B, s);
        A.dispatch([&](auto nd) {
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) f(i, A.template row<nd()>(i, xv));
        });
    } else {
        tile_spmv<double>(r0, r1, B.rp, B.ci, B.v64, xv, prod, f);
    }
}

// K1 rows [r0, r1) of a CSR system whose tile holds at most kLdsNnz nonzeros
// (k0 = rp[r0], k1 = rp[r1]):
// p_new = D^-1 r + beta p_old for the tile's own rows (stored to pn) and for
// every column its nonzeros touch (the same expression on the same inputs as
// the column's owner row, so bit-identical), q = A p_new (stored) and the
// p_new . q partial.  Every column and value load of the tile is issued, then
// every gather of r, p_old (and D^-1) at those columns, then the own rows'
// loads, and only then is anything stored or summed: a wave keeps its loads in
// flight together instead of one dependent gather chain per product.  The
// products val * p_new go to LDS and each row is summed from there exactly as
// tile_spmv sums it (same order, same rounding).  The trip count nu is
// uniform over the workgroup; lanes past the last nonzero re-read it and park
// a product in a slot no row reads.
constexpr int kCsrOwn = 4;  // own rows per thread held in registers (tiles of up to 1024 rows)
template <class T>
__device__ __forceinline__ void k1_csr(int r0, int r1, const int* __restrict__ rp, const int* __restrict__ ci,
                                       const T* __restrict__ val, const T* __restrict__ r, const T* __restrict__ po,
                                       const T* __restrict__ dinv, bool cdok, T cdv, double beta, lds_ptr<T> prod,
                                       T* __restrict__ pn, T* __restrict__ q, int k0, int k1, double& pq) {
    // products per thread, in rounds of UB whose loads are all in flight
    // together (FP64: two rounds of four, which keeps the kernel's registers
    // at the occupancy of its other paths)
    constexpr int U = kLdsNnz / kBlock, UB = sizeof(T) == 8 ? 4 : U;
    const int tid = threadIdx.x;
    const int nu = (k1 - k0 + kBlock - 1) / kBlock;
    auto pnew = [&](T d, T rj, T pj) -> T { return (T)((double)(T)(d * rj) + beta * (double)pj); };
    T ro[kCsrOwn], pw[kCsrOwn], dw[kCsrOwn];
#pragma unroll
    for (int ub = 0; ub < U; ub += UB) {
        unsigned col[UB];
        T vv[UB], rg[UB], pg[UB], dg[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            col[u] = 0;
            vv[u] = rg[u] = pg[u] = T(0);
            dg[u] = cdv;
        }
#pragma unroll
        for (int u = 0; u < UB; ++u)
            if (ub + u < nu) {
                const int k = min(k0 + tid + (ub + u) * kBlock, k1 - 1);
                col[u] = (unsigned)ci[k];
                vv[u] = val[k];
            }
#pragma unroll
        for (int u = 0; u < UB; ++u)
            if (ub + u < nu) {
                rg[u] = r[col[u]];
                pg[u] = po[col[u]];
                if (!cdok) dg[u] = dinv[col[u]];
            }
        if (ub == 0) {
#pragma unroll
            for (int j = 0; j < kCsrOwn; ++j) {
                const int i = min(r0 + tid + j * kBlock, r1 - 1);
                ro[j] = r[i];
                pw[j] = po[i];
                dw[j] = cdok ? cdv : dinv[i];
            }
        }
#pragma unroll
        for (int u = 0; u < UB; ++u)
            if (ub + u < nu) prod[tid + (ub + u) * kBlock] = vv[u] * pnew(dg[u], rg[u], pg[u]);
    }
    T own[kCsrOwn];
#pragma unroll
    for (int j = 0; j < kCsrOwn; ++j) {
        const int i = r0 + tid + j * kBlock;
        own[j] = pnew(dw[j], ro[j], pw[j]);
        if (i < r1) pn[i] = own[j];
    }
    auto pnew_at = [&](int i) -> T { return pnew(cdok ? cdv : dinv[i], r[i], po[i]); };
    for (int i = r0 + tid + kCsrOwn * kBlock; i < r1; i += kBlock) pn[i] = pnew_at(i);
    __syncthreads();
    const int nrows = r1 - r0;
    int tpr = 1;
    while (tpr < 64 && tpr * 2 * nrows <= kBlock) tpr *= 2;
    auto row_sum = [&](int row, int l) -> T {
        const int a = rp[row] - k0, b = rp[row + 1] - k0;
        T s = 0;
        for (int k = a + l; k < b; k += tpr) s += prod[k];
        return s;
    };
    if (tpr == 1) {  // one row per thread: rows r0 + tid + j kBlock, p_new from registers
#pragma unroll
        for (int j = 0; j < kCsrOwn; ++j) {
            const int row = r0 + tid + j * kBlock;
            if (row < r1) {
                const T sum = row_sum(row, 0);
                q[row] = sum;
                pq += (double)own[j] * (double)sum;
            }
        }
        for (int row = r0 + tid + kCsrOwn * kBlock; row < r1; row += kBlock) {
            const T sum = row_sum(row, 0);
            q[row] = sum;
            pq += (double)pnew_at(row) * (double)sum;
        }
    } else {
        const int g = tid / tpr, l = tid - g * tpr, groups = kBlock / tpr;
        for (int row = r0 + g; row < r1; row += groups) {
            T s = row_sum(row, l);
            for (int o = tpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (l == 0) {
                q[row] = s;
                pq += (double)pnew_at(row) * (double)s;
            }
        }
    }
}

// K2 rows [r0, r1) of a system whose q is stored (CSR, DIA): x += alpha p
// (x_lo in mixed mode), r -= alpha q, r.r and r.z partials.  kCsrOwn rows per
// thread load together before any store (vmcnt counts loads and stores in one
// queue), each row in the same arithmetic as k_b_k2's row body.
template <class T, bool MIXED>
__device__ __forceinline__ void k2_rows(int r0, int r1, const T* __restrict__ p, const T* __restrict__ q,
                                        T* __restrict__ r, double* __restrict__ x64, float* __restrict__ xlo,
                                        const T* __restrict__ dinv, bool cdok, T cdv, double alpha,
                                        double (&red)[2]) {
    using X = std::conditional_t<MIXED, float, double>;
    X* __restrict__ xs = MIXED ? (X*)xlo : (X*)x64;
    for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += kCsrOwn * kBlock) {
        T pv[kCsrOwn], qv[kCsrOwn], rv[kCsrOwn], dv[kCsrOwn];
        X xv[kCsrOwn];
#pragma unroll
        for (int j = 0; j < kCsrOwn; ++j) {
            const int i = min(i0 + j * kBlock, r1 - 1);
            pv[j] = p[i];
            qv[j] = q[i];
            rv[j] = r[i];
            xv[j] = xs[i];
            dv[j] = cdok ? cdv : dinv[i];
        }
#pragma unroll
        for (int j = 0; j < kCsrOwn; ++j) {
            const int i = i0 + j * kBlock;
            if (i < r1) {
                xs[i] = (X)((double)xv[j] + alpha * (double)pv[j]);
                const T ri = (T)((double)rv[j] - alpha * (double)qv[j]);
                r[i] = ri;
                const double rd = (double)ri;
                red[0] += rd * rd;
                red[1] += rd * (double)(T)(dv[j] * ri);
            }
        }
    }
}

// Per-system finish: one wave per system of [k0, k0+gridDim.x), or per entry
// of sys when given (the active systems of an iteration).  The
// system's state and tile range are loaded first (one round trip), then its
// tile partials; op runs on a register copy of the state in lane 0, which
// writes it back.  The partials are summed in the order of a 256-thread
// block_sum: "thread" 64 w + l (w = 0..3) adds tiles t0 + 64 w + l + 256 j,
// each group of 64 is summed by the xor butterfly of wave_sum, and the four
// group sums are added in order, so one wave gives the same bits.
constexpr int kFinishThreads = 64;
template <int NV, class Op>
__global__ void __launch_bounds__(kFinishThreads) k_b_finish(BArgs B, int k0, const int* __restrict__ sys, Op op) {
    const int s = sys ? sys[blockIdx.x] : k0 + blockIdx.x;
    const int t0 = B.sys_t0[s], t1 = B.sys_t1[s];
    BState S = B.st[s];
    if (!op.active(S)) return;
    const int lane = threadIdx.x;
    constexpr int W = kBlock / 64;
    // U steps of every group's chain are loaded before they are added (a
    // system of 2048 tiles is 8 steps per chain); a slot past t1 adds +0.0,
    // which leaves a sum that starts at +0.0 unchanged
    constexpr int U = 4;
    double acc[W][NV];
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[w][k] = 0.0;
    for (int j0 = t0 + lane; j0 < t1; j0 += U * kBlock) {
        double v[U][W][NV];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const int t = j0 + u * kBlock + 64 * w;
#pragma unroll
                for (int k = 0; k < NV; ++k) v[u][w][k] = t < t1 ? B.partials[(size_t)k * B.T + t] : 0.0;
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int w = 0; w < W; ++w)
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[w][k] += v[u][w][k];
    }
    double tot[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) sum += wave_sum(acc[w][k]);
        tot[k] = sum;
    }
    if (lane == 0) {
        op(S, tot, B);
        B.st[s] = S;
    }
}

template <int NV, int BS = kBlock>
__device__ __forceinline__ void tile_partial(double (&v)[NV], const BArgs& B, int tile) {
    __shared__ double red[NV * (BS / 64)];
    block_sum<NV, BS>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) B.partials[(size_t)k * B.T + tile] = v[k];
    }
}

// XCD-aware tile order.  Workgroup b of a launch runs on XCD b % 8 (the
// dispatcher deals workgroups round-robin over the eight XCDs, each with its
// own L2).  XCD x takes runs of kXcdRun consecutive tiles (x, x+8, x+16, ...
// runs), so neighbouring tiles, whose stencil halos overlap, mostly share an
// L2, while every large system still spreads over all XCDs (one contiguous
// range per XCD would leave XCDs idle as systems converge).  A bijection of
// [0, nt): the last nt mod (8 kXcdRun) tiles keep the plain order.
// Runs of 8 measured best on MI355X (512-system stencil batch, 2048-row
// tiles: 2.6% faster than plain order; runs of 2, 4 and 16 in between).
constexpr int kXcds = 8, kXcdRun = 8;
__device__ __forceinline__ int xcd_tile(int b, int nt) {
    constexpr int span = kXcds * kXcdRun;
    if (b >= nt / span * span) return b;
    const int x = b % kXcds, slot = b / kXcds;
    return ((slot / kXcdRun) * kXcds + x) * kXcdRun + slot % kXcdRun;
}

// ---- init: r64 = b - A x ; |r|, |b| --------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_b_init1(BArgs B, int t0, const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r64) {
    __shared__ double prod[kLdsNnz];
    const int tile = t0 + blockIdx.x;
    const int r0 = B.tile_r0[tile], r1 = B.tile_r1[tile];
    double red[2] = {0.0, 0.0};
    tile_rows64(B, B.tile_sys[tile], r0, r1, [&](int j) { return x[j]; }, prod, [&](int i, double sum) {
        const double ri = b[i] - sum;
        r64[i] = ri;
        red[0] += ri * ri;
    });
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) red[1] += b[i] * b[i];
    tile_partial<2>(red, B, tile);
    if (threadIdx.x < 8) {  // the tile's tagged partials of both phases start untagged
        const int ph = threadIdx.x >> 2, k = threadIdx.x & 3;
        B.tagp[((size_t)ph * B.T + tile) * 4 + k] = 0ULL;
    }
}

struct OpInit1 {
    __device__ bool active(const BState&) const { return true; }
    __device__ void operator()(BState& S0, const double* tot, const BArgs& B) const {
        BState S{};
        S.bnorm = sqrt(tot[1]);
        const double r0n = sqrt(tot[0]);
        S.rmax = r0n;
        S.r0 = r0n;
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

// r = (T) r64, x_lo = 0, p = D^-1 r (first p buffer), r.z partials
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
        S.beta = 0.0;  // the first K1 forms p = z + 0 * z
        S.pcur = 0;
    }
};

// ---- the per-system decisions after K1 (F1) and K2 (F2) -------------------------------------
struct OpF1 {
    __device__ bool active(const BState& S) const { return !S.halt; }
    __device__ void operator()(BState& S, const double* tot, const BArgs& B) const {
        if (S.flag != BS_FLUSH) {
            const double pq = tot[0];
            if (fabs(pq) < kEps * 1e6 * fabs(S.rho)) {  // cg_solver.cpp:109-120: restart
                S.iter += 1;
                S.restarts += 1;
                S.restarts_pq += 1;
                S.flag = BS_FLUSH;
                S.reason = BR_RESTART;
                if (S.iter >= B.max_iter) S.halt = 1;
            } else {
                S.alpha = S.rho / pq;
                S.flag = BS_NORMAL;
                S.pcur ^= 1;  // K1 wrote p_new into the other buffer
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
        if (S.iter >= B.max_iter) {  // a check or reliable update on the last allowed iteration
            S.halt = 1;
            S.flag = BS_NORMAL;
            S.reason = BR_NONE;
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
        S.flag = BS_FLUSHED;
        S.reason = BR_NONE;
    }
};

template <bool MIXED>
struct OpF2 {
    __device__ bool active(const BState& S) const { return !S.halt && S.flag == BS_NORMAL; }
    __device__ void operator()(BState& S, const double* tot, const BArgs& B) const {
        const double rn = sqrt(tot[0]);
        S.rel = rn / S.bnorm;
        S.iter += 1;
        if (S.rel < B.tol) {  // cg_solver.cpp:143, confirmed on the FP64 true residual
            S.flag = BS_FLUSH;
            S.reason = BR_CHECK;
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

// ---- finishes inside the iteration kernels -------------------------------------------------
// An iteration kernel's workgroup publishes its tile's partial sums as tagged
// 64-bit words (tag << 32 | 32 bits of the value, two words a value, written
// with relaxed device-scope atomics: each word is valid exactly when its tag
// is the launch's, so no flag, fence or cache writeback orders anything).  One
// workgroup per system — the one the dispatcher starts last (blocks start in
// increasing index order; its record carries fin = 1) — then waits for the
// tags of the system's other tiles, which all started before it and wait on
// nothing, and runs the system's finish: the sum of k_b_finish in the same
// order, then the decisions.  No atomics and no extra launch.  The tag is the
// system's finish count + 1, read with the state at the kernel's start, so
// every launch tags differently; a system's tags are cleared when it starts.
// A wait that never sees its tag halts the system after 2 s (BR_STALLED)
// instead of hanging.
constexpr long long kFinishWaitTicks = 200000000LL;  // 2 s of s_memrealtime (100 MHz)

__device__ __forceinline__ void tag_st(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long tag_ld(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV, int BS, class Op>
__device__ __forceinline__ void tile_finish(double (&v)[NV], const BArgs& B, int tile, int s, int phase,
                                            unsigned tag, int fin, Op op) {
    static_assert(BS >= kBlock && BS % kBlock == 0, "the finish uses the first 256 threads");
    __shared__ double red[NV * (BS / 64)];
    block_sum<NV, BS>(v, red);
    if (threadIdx.x == 0) {
        unsigned long long* w = B.tagp + ((size_t)phase * B.T + tile) * (2 * NV);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(v[k]);
            tag_st(w + 2 * k, ((unsigned long long)tag << 32) | (bits & 0xffffffffULL));
            tag_st(w + 2 * k + 1, ((unsigned long long)tag << 32) | (bits >> 32));
        }
    }
    if (!fin) return;  // uniform over the workgroup
    __syncthreads();   // red[] is reused below
    // k_b_finish's sum over the tagged words, as a 256-thread tree: thread
    // 64 w + l adds tiles t0 + 64 w + l + 256 j in order, each wave's sums go
    // through the xor butterfly, the four wave sums are added in order
    const int tid = threadIdx.x;
    const int t0 = B.sys_t0[s], t1 = B.sys_t1[s];
    constexpr int U = 2;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    const long long t_start = (long long)__builtin_amdgcn_s_memrealtime();
    int stalled = 0;
    for (int j0 = t0 + tid; tid < kBlock && j0 < t1; j0 += U * kBlock) {  // threads >= 256 add nothing
        unsigned long long wd[U][NV][2];
        for (;;) {
            bool ok = true;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = j0 + u * kBlock;
                const unsigned long long* q = B.tagp + ((size_t)phase * B.T + t) * (2 * NV);
#pragma unroll
                for (int k = 0; k < NV; ++k)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        wd[u][k][h] = t < t1 ? tag_ld(q + 2 * k + h) : ((unsigned long long)tag << 32);
                        ok = ok && (unsigned)(wd[u][k][h] >> 32) == tag;
                    }
            }
            if (__builtin_amdgcn_ballot_w64(!ok) == 0) break;
            if ((long long)__builtin_amdgcn_s_memrealtime() - t_start > kFinishWaitTicks) {
                stalled = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (stalled) break;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < NV; ++k)
                acc[k] += __longlong_as_double((long long)((wd[u][k][0] & 0xffffffffULL) | (wd[u][k][1] << 32)));
    }
    __shared__ double wsum[NV * (kBlock / 64)];
    __shared__ int wstall[kBlock / 64];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = wave_sum(acc[k]);
    if ((tid & 63) == 0 && tid < kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) wsum[k * (kBlock / 64) + (tid >> 6)] = acc[k];
        wstall[tid >> 6] = stalled;
    }
    __syncthreads();
    if (tid == 0) {
        BState S = B.st[s];
        double tot[NV];
        int st = 0;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double sum = 0.0;
#pragma unroll
            for (int ww = 0; ww < kBlock / 64; ++ww) sum += wsum[k * (kBlock / 64) + ww];
            tot[k] = sum;
        }
#pragma unroll
        for (int ww = 0; ww < kBlock / 64; ++ww) st |= wstall[ww];
        if (st) {
            S.halt = 1;
            S.reason = BR_STALLED;
        } else {
            op(S, tot, B);
        }
        S.seq += 1;
        B.st[s] = S;
    }
}

// ---- K1 / F1 -------------------------------------------------------------------------------
// NORMAL / FLUSHED: p_new = D^-1 r + beta p_old into the other p buffer, and
// q = A p_new with p.q partials.  The p update of cg_solver.cpp:173-176 is
// evaluated on the fly for every column the tile's rows touch (the same
// expression on the same inputs as the row's owner, so bit-identical), which
// removes a pass over r and p per iteration.  Stencil-stored systems do not
// store q: K2 recomputes it from p_new.
// FLUSH: FP64 true residual r = b - A (x64 + x_lo).
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k1(BArgs B, T* __restrict__ p0, T* __restrict__ p1, T* __restrict__ q, T* __restrict__ r,
       const double* __restrict__ b, const double* __restrict__ x64, const float* __restrict__ xlo) {
    extern __shared__ __attribute__((aligned(16))) double stage_lds[];  // B.stage_doubles
    // every access to the staging buffer goes through address-space-3 pointers
    const lds_ptr<double> prod64 = (lds_ptr<double>)stage_lds;
    const TileVec<T>& D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    if (D.vec) return;  // k_b_k1v runs it
    const int tile = D.tile;
    const int s = B.tile_sys[tile];  // (D.sys, the same, puts a stack frame on this kernel)
    const BState* S0 = &B.st[s];
    if (S0->halt) return;
    const int flag = S0->flag;
    const unsigned tag = (unsigned)S0->seq + 1u;
    const int r0 = D.r0, r1 = D.r1;
    const bool cdok = B.cd_ok[s] != 0;
    const T cdv = Prec<T>::cd(B)[s];
    const T* __restrict__ dinv = Prec<T>::dinv(B);
    double red[2] = {0.0, 0.0};
    if (flag == BS_FLUSH) {
        auto xv = [&](int j) { return MIXED ? x64[j] + (double)xlo[j] : x64[j]; };
        tile_rows64(B, s, r0, r1, xv, prod64, [&](int i, double sum) {
            const double rh = b[i] - sum;
            const T ri = (T)rh;
            r[i] = ri;
            red[0] += rh * rh;
            red[1] += (double)ri * (double)(T)((cdok ? cdv : dinv[i]) * ri);
        });
        tile_finish<2, kBlock>(red, B, tile, s, 0, tag, D.fin, OpF1{});
        return;
    }
    const double beta = S0->beta;
    const T* __restrict__ po = S0->pcur ? p1 : p0;
    T* __restrict__ pn = S0->pcur ? p0 : p1;
    auto pnew = [&](int j) -> T {
        const T dj = cdok ? cdv : dinv[j];
        return (T)((double)(T)(dj * r[j]) + beta * (double)po[j]);
    };
    const int kind = B.dia_kind[s];
    if (kind != SK_CSR) {
        const Stencil<T> A(B, s);
        const lds_ptr<T> lds = (lds_ptr<T>)prod64;
        const int cap = B.stage_doubles * (int)(sizeof(double) / sizeof(T));
        auto out = [&](int i, T pi, T sum) {
            if (kind == SK_DIA) q[i] = sum;
            red[0] += (double)pi * (double)sum;
        };
        if (stage_need(A.halo, (int)sizeof(T)) <= cap) {  // p_new staged in LDS, evaluated once per row
            A.dispatch([&](auto nd) {
                stencil_staged<nd()>(A, r0, r1, lds, cap, pnew, [&](int j, T v) { pn[j] = v; }, out);
            });
        } else {  // halo wider than the staging buffer: rows straight from memory
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
                const T pi = pnew(i);
                pn[i] = pi;
                out(i, pi, A.template row<0>(i, pnew));
            }
        }
    } else {
        const int k0 = B.rp[r0], k1 = B.rp[r1];
        if (k1 - k0 <= kLdsNnz) {
            k1_csr<T>(r0, r1, B.rp, B.ci, Prec<T>::val(B), r, po, dinv, cdok, cdv, beta, (lds_ptr<T>)prod64, pn, q,
                      k0, k1, red[0]);
        } else {  // one row longer than the product stage
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) pn[i] = pnew(i);
            tile_spmv<T>(r0, r1, B.rp, B.ci, Prec<T>::val(B), pnew, (lds_ptr<T>)prod64, [&](int i, T sum) {
                q[i] = sum;
                red[0] += (double)pnew(i) * (double)sum;
            });
        }
    }
    tile_finish<2, kBlock>(red, B, tile, s, 0, tag, D.fin, OpF1{});
}

// ---- K2 / F2 -------------------------------------------------------------------------------
// NORMAL: x += alpha p, r -= alpha q (q recomputed from p for stencil-stored
// systems), r.r / r.z partials.  FLUSHED: x64 += x_lo, x_lo = 0.
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k2(BArgs B, const T* __restrict__ p0, const T* __restrict__ p1, const T* __restrict__ q,
       T* __restrict__ r, double* __restrict__ x64, float* __restrict__ xlo) {
    extern __shared__ __attribute__((aligned(16))) double stage_lds[];  // B.stage_doubles
    const TileVec<T>& D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    if (D.vec) return;  // k_b_k2v runs it
    const int tile = D.tile;
    const int s = D.sys;
    const BState* S0 = &B.st[s];
    if (S0->halt) return;
    const int flag = S0->flag;
    const unsigned tag = (unsigned)S0->seq + 1u;
    const int r0 = D.r0, r1 = D.r1;
    if (flag == BS_FLUSHED) {
        if (MIXED)
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        return;
    }
    if (flag != BS_NORMAL) return;
    const double alpha = S0->alpha;
    const T* __restrict__ p = S0->pcur ? p1 : p0;
    const bool cdok = B.cd_ok[s] != 0;
    const T cdv = Prec<T>::cd(B)[s];
    const T* __restrict__ dinv = Prec<T>::dinv(B);
    double red[2] = {0.0, 0.0};
    auto body = [&](int i, T pi, T qi) {
        if (MIXED) xlo[i] = (float)((double)xlo[i] + alpha * (double)pi);
        else x64[i] = x64[i] + alpha * (double)pi;
        const T ri = (T)((double)r[i] - alpha * (double)qi);
        r[i] = ri;
        const double rd = (double)ri;
        red[0] += rd * rd;
        red[1] += rd * (double)(T)((cdok ? cdv : dinv[i]) * ri);
    };
    if (B.dia_kind[s] == SK_STENCIL) {
        const Stencil<T> A(B, s);
        const lds_ptr<T> lds = (lds_ptr<T>)(lds_ptr<double>)stage_lds;
        const int cap = B.stage_doubles * (int)(sizeof(double) / sizeof(T));
        auto pv = [&](int j) { return p[j]; };
        if (stage_need(A.halo, (int)sizeof(T)) <= cap) {
            A.dispatch([&](auto nd) { stencil_staged<nd()>(A, r0, r1, lds, cap, pv, [](int, T) {}, body); });
        } else {
            for (int i = r0 + threadIdx.x; i < r1; i += kBlock) body(i, p[i], A.template row<0>(i, pv));
        }
    } else {
        k2_rows<T, MIXED>(r0, r1, p, q, r, x64, xlo, dinv, cdok, cdv, alpha, red);
    }
    tile_finish<2, kBlock>(red, B, tile, s, 1, tag, D.fin, OpF2<MIXED>{});
}

// ---- K1 / K2 of the 16-byte five-point path ------------------------------------------------
// Tiles whose record has vec = 1 (BS = kVecNarrow) or vec = 2 (BS = kVecWide)
// run here in every state (the general kernels skip them).  Kernels of their own keep the general paths' register
// budget out of these, and every value they need before the first vector load
// comes from the tile record and the system's state.
// K1: NORMAL / FLUSHED as k_b_k1's stencil path; FLUSH: the FP64 true
// residual r = b - A (x64 + x_lo) with the rows read straight from memory
// (Stencil::row<5>, the same sum as k_b_k1's tile_rows64).
template <class T, bool MIXED, int BS>
__global__ void __launch_bounds__(BS, 8)
k_b_k1v(BArgs B, T* __restrict__ p0, T* __restrict__ p1, T* __restrict__ r, const double* __restrict__ b,
        const double* __restrict__ x64, const float* __restrict__ xlo) {
    extern __shared__ __attribute__((aligned(16))) double stage_lds[];  // B.stage_doubles
    const TileVec<T> D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    const int tile = D.tile;
    if (D.vec != (BS == kVecNarrow ? 1 : 2)) return;
    const BHot h = load_hot(&B.st[D.sys]);
    const int mode = h.halt ? -1 : h.flag;  // one branch on both fields: their loads issue together
    if (mode < 0) return;
    double red[2] = {0.0, 0.0};
    if (mode == BS_FLUSH) {
        const Stencil<double> A(B, D.sys);
        const T dv = D.cd;
        auto xv = [&](int j) { return MIXED ? x64[j] + (double)xlo[j] : x64[j]; };
        for (int i = D.r0 + threadIdx.x; i < D.r1; i += BS) {
            const double rh = b[i] - A.template row<5>(i, xv);
            const T ri = (T)rh;
            r[i] = ri;
            red[0] += rh * rh;
            red[1] += (double)ri * (double)(T)(dv * ri);
        }
    } else {
        k1_stencil_vec<T, BS>(D, B.umask, (lds_ptr<T>)(lds_ptr<double>)stage_lds, r, h.pcur ? p1 : p0, h.pcur ? p0 : p1,
                       h.beta, red[0]);
    }
    tile_finish<2, BS>(red, B, tile, D.sys, 0, (unsigned)h.seq + 1u, D.fin, OpF1{});
}

// K2: NORMAL as k_b_k2's stencil path; FLUSHED: x64 += x_lo, x_lo = 0.
template <class T, bool MIXED, int BS>
__global__ void __launch_bounds__(BS, sizeof(T) == 4 ? 8 : 7)
k_b_k2v(BArgs B, const T* __restrict__ p0, const T* __restrict__ p1, T* __restrict__ r,
        double* __restrict__ x64, float* __restrict__ xlo) {
    extern __shared__ __attribute__((aligned(16))) double stage_lds[];  // B.stage_doubles
    const TileVec<T> D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    const int tile = D.tile;
    if (D.vec != (BS == kVecNarrow ? 1 : 2)) return;
    const BHot h = load_hot(&B.st[D.sys]);
    const int mode = h.halt ? -1 : h.flag;  // one branch on both fields: their loads issue together
    if (mode != BS_NORMAL && mode != BS_FLUSHED) return;
    if (mode == BS_FLUSHED) {
        if (MIXED)
            for (int i = D.r0 + threadIdx.x; i < D.r1; i += BS) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        return;
    }
    double red[2] = {0.0, 0.0};
    k2_stencil_vec<T, MIXED, BS>(D, B.umask, (lds_ptr<T>)(lds_ptr<double>)stage_lds, h.pcur ? p1 : p0, r, x64, xlo,
                             h.alpha, red);
    tile_finish<2, BS>(red, B, tile, D.sys, 1, (unsigned)h.seq + 1u, D.fin, OpF2<MIXED>{});
}

// ---- K1 / K2 of CSR systems (TileVec::vec = 3) ---------------------------------------------
// The same steps and arithmetic as k_b_k1 / k_b_k2's CSR paths, with every
// value a workgroup needs before its first vector load in the tile record and
// the loads that do not depend on the system's state (the tile's nonzero
// range; K2's own rows of q, r and x) issued together with the state: a
// workgroup waits for the record, then one round trip, then its column and
// value loads, where the general kernel waited for four dependent lookups.
template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k1c(BArgs B, T* __restrict__ p0, T* __restrict__ p1, T* __restrict__ q, T* __restrict__ r,
        const double* __restrict__ b, const double* __restrict__ x64, const float* __restrict__ xlo) {
    extern __shared__ __attribute__((aligned(16))) double stage_lds[];  // B.stage_doubles >= kLdsNnz
    const TileVec<T> D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    if (D.vec != 3) return;
    const int k0 = B.rp[D.r0], k1 = B.rp[D.r1];
    const BHot h = load_hot(&B.st[D.sys]);
    const int mode = h.halt ? -1 : h.flag;
    if (mode < 0) return;
    const T* __restrict__ dinv = Prec<T>::dinv(B);
    const bool cdok = D.cdok != 0;
    double red[2] = {0.0, 0.0};
    if (mode == BS_FLUSH) {  // FP64 true residual, as k_b_k1
        auto xv = [&](int j) { return MIXED ? x64[j] + (double)xlo[j] : x64[j]; };
        tile_spmv<double>(D.r0, D.r1, B.rp, B.ci, B.v64, xv, (lds_ptr<double>)stage_lds, [&](int i, double sum) {
            const double rh = b[i] - sum;
            const T ri = (T)rh;
            r[i] = ri;
            red[0] += rh * rh;
            red[1] += (double)ri * (double)(T)((cdok ? D.cd : dinv[i]) * ri);
        });
    } else {
        k1_csr<T>(D.r0, D.r1, B.rp, B.ci, Prec<T>::val(B), r, h.pcur ? p1 : p0, dinv, cdok, D.cd, h.beta,
                  (lds_ptr<T>)(lds_ptr<double>)stage_lds, h.pcur ? p0 : p1, q, k0, k1, red[0]);
    }
    tile_finish<2, kBlock>(red, B, D.tile, D.sys, 0, (unsigned)h.seq + 1u, D.fin, OpF1{});
}

template <class T, bool MIXED>
__global__ void __launch_bounds__(kBlock)
k_b_k2c(BArgs B, const T* __restrict__ p0, const T* __restrict__ p1, const T* __restrict__ q,
        T* __restrict__ r, double* __restrict__ x64, float* __restrict__ xlo) {
    using X = std::conditional_t<MIXED, float, double>;
    const TileVec<T> D = Prec<T>::tva(B)[xcd_tile(blockIdx.x, gridDim.x)];
    if (D.vec != 3) return;
    X* __restrict__ xs = MIXED ? (X*)xlo : (X*)x64;
    const int tid = threadIdx.x, r0 = D.r0, r1 = D.r1;
    // the first kCsrOwn rows per thread of q, r and x, loaded with the state
    T qv[kCsrOwn], rv[kCsrOwn];
    X xv[kCsrOwn];
#pragma unroll
    for (int j = 0; j < kCsrOwn; ++j) {
        const int i = min(r0 + tid + j * kBlock, r1 - 1);
        qv[j] = q[i];
        rv[j] = r[i];
        xv[j] = xs[i];
    }
    const BHot h = load_hot(&B.st[D.sys]);
    const int mode = h.halt ? -1 : h.flag;
    if (mode != BS_NORMAL && mode != BS_FLUSHED) return;
    if (mode == BS_FLUSHED) {
        if (MIXED)
            for (int i = r0 + tid; i < r1; i += kBlock) {
                x64[i] += (double)xlo[i];
                xlo[i] = 0.0f;
            }
        return;
    }
    const T* __restrict__ p = h.pcur ? p1 : p0;
    const T* __restrict__ dinv = Prec<T>::dinv(B);
    const bool cdok = D.cdok != 0;
    const double alpha = h.alpha;
    double red[2] = {0.0, 0.0};
    T pv[kCsrOwn], dv[kCsrOwn];
#pragma unroll
    for (int j = 0; j < kCsrOwn; ++j) {
        const int i = min(r0 + tid + j * kBlock, r1 - 1);
        pv[j] = p[i];
        dv[j] = cdok ? D.cd : dinv[i];
    }
    // k2_rows' row body, rows r0 + tid + j kBlock in the same order
#pragma unroll
    for (int j = 0; j < kCsrOwn; ++j) {
        const int i = r0 + tid + j * kBlock;
        if (i < r1) {
            xs[i] = (X)((double)xv[j] + alpha * (double)pv[j]);
            const T ri = (T)((double)rv[j] - alpha * (double)qv[j]);
            r[i] = ri;
            const double rd = (double)ri;
            red[0] += rd * rd;
            red[1] += rd * (double)(T)(dv[j] * ri);
        }
    }
    if (r1 - r0 > kCsrOwn * kBlock)
        k2_rows<T, MIXED>(r0 + kCsrOwn * kBlock, r1, p, q, r, x64, xlo, dinv, cdok, D.cd, alpha, red);
    tile_finish<2, kBlock>(red, B, D.tile, D.sys, 1, (unsigned)h.seq + 1u, D.fin, OpF2<MIXED>{});
}

// ---- final: x64 += x_lo and the FP64 true residual per (large) system ----------------------
__global__ void __launch_bounds__(kBlock)
k_b_flush_tiles(BArgs B, int t0, double* __restrict__ x64, float* __restrict__ xlo) {
    const int r0 = B.tile_r0[t0 + blockIdx.x], r1 = B.tile_r1[t0 + blockIdx.x];
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
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
    tile_rows64(B, B.tile_sys[tile], r0, r1, [&](int j) { return x[j]; }, prod, [&](int i, double sum) {
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
    const bool is_dia = B.dia_kind[s] != SK_CSR;
    const Stencil<T> A(B, s);
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
    S.r0 = r0n;
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
            if (is_dia) {
                A.dispatch([&](auto nd) {
                    for (int i = a + tid; i < e; i += kSmallThreads) {
                        const T sum = A.template row<nd()>(i, [&](int j) { return p[j]; });
                        q[i] = sum;
                        v1[0] += (double)p[i] * (double)sum;
                    }
                });
            } else {
                for (int i = a + tid; i < e; i += kSmallThreads) {
                    T sum = 0;
                    for (int k = rp[i]; k < rp[i + 1]; ++k) sum += val[k] * p[ci[k]];
                    q[i] = sum;
                    v1[0] += (double)p[i] * (double)sum;
                }
            }
            wg_sum<1>(v1, red);
            int reason = BR_NONE;
            const double pq = v1[0];
            if (fabs(pq) < kEps * 1e6 * fabs(S.rho)) {
                S.iter += 1;
                S.restarts += 1;
                S.restarts_pq += 1;
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
                if (S.rel < B.tol) {  // confirmed on the FP64 true residual below
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
    int capture_failures = 0;  // chunk captures that failed and ran as direct launches
    std::vector<int> h_row_off;  // user order, S+1
    std::vector<int> h_rp;       // global row_ptr, N+1
    std::vector<int> order;      // internal position -> user system id (by nnz, descending)
    DevBuf rp, ci, v64, v32, dinv64, dinv32, sys_row0, sys_row1;
    DevBuf dia_kind, dia_nd, dia_off, dia_base, dia64, dia32, uv64, uv32, umask, cd_ok, cd64, cd32;
    int n_kind[3] = {0, 0, 0};  // systems per SK_* storage
    std::vector<int> h_kind;    // SK_* per internal system
    // host copies of the storage tables (per internal system) for the tile records
    std::vector<int> h_nd, h_off, h_cdok;
    std::vector<double> h_uv, h_cd;
    int stage_doubles[2] = {kStageMinDoubles, kStageMinDoubles};  // K1 / K2 LDS: [0] FP64, [1] FP32
    // tiled schedule: rows per tile for CSR systems (0 = not tuned yet; also
    // capped at kLdsNnz nonzeros) and for diagonal-stored systems
    int tile_rows = 0, tile_rows_dia = 0;
    int T = 0;
    DevBuf tile_r0, tile_r1, tile_sys, sys_t0, sys_t1, tv32, tv64;
    // active-tile compaction: host copies of the records, the compacted device
    // copies K1 / K2 read, and two pinned staging buffers (an upload never
    // overwrites the one still being copied)
    std::vector<TileVec<float>> h_tv32;
    std::vector<TileVec<double>> h_tv64;
    DevBuf tva32, tva64;
    void* pin_act[2] = {nullptr, nullptr};
    size_t pin_act_bytes = 0;
    int pin_next = 0;
    std::vector<int> h_sys_t0, h_sys_t1;
    std::vector<std::pair<int, int>> waves;  // internal system ranges within [0, n_large)
    // per-solve work (reused)
    DevBuf st, partials, r64, r, p, p1, q, xlo;
    DevBuf tagp;  // tagged partials of the iteration kernels
    int work_prec = -1;
    // a captured chunk replays the kernel arguments it was captured with
    // (tolerance, iteration cap, tiles, ...), so the whole argument block is
    // part of its key
    struct G {
        hipGraphExec_t exec = nullptr;
        int wave = -1, prec = -1, chunk = 0, nt = 0, ns = 0, kinds = 0;
        BArgs args{};
    };
    std::vector<G> graphs;
    const double* graph_b = nullptr;  // the b and x the captured kernels point at
    const double* graph_x = nullptr;
    ~Batch() {
        drop_graphs();
        for (void* q : pin_act)
            if (q) (void)hipHostFree(q);
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

// One TileVec<T> per tile.  vec is set where the 16-byte five-point path
// applies: a stencil-stored system with at most five diagonals, a scalar
// D^-1, and a chunk and halo that fit the precision's LDS window.
template <class T>
static void build_tile_records(Batch& B, const std::vector<int>& r0, const std::vector<int>& r1,
                               const std::vector<int>& sys, DevBuf& out) {
    const int es = (int)sizeof(T), cap = B.stage_doubles[es == 4 ? 1 : 0] * 8 / es;
    std::vector<TileVec<T>> tv(std::max<size_t>(r0.size(), 1));
    std::memset(tv.data(), 0, tv.size() * sizeof(TileVec<T>));
    // CSR systems all of whose tiles fit the product stage (no row longer than it)
    std::vector<char> csr_fits(B.S, 1);
    for (size_t t = 0; t < r0.size(); ++t)
        if (B.h_rp[r1[t]] - B.h_rp[r0[t]] > kLdsNnz) csr_fits[sys[t]] = 0;
    for (size_t t = 0; t < r0.size(); ++t) {
        const int k = sys[t], s = B.order[k];
        TileVec<T>& d = tv[t];
        d.r0 = r0[t];
        d.r1 = r1[t];
        d.sys = k;
        d.tile = (int)t;
        d.row0 = B.h_row_off[s];
        d.last = B.h_row_off[s + 1] - 1;
        const int nd = B.h_nd[k];
        int halo = 0;
        for (int j = 0; j < nd; ++j) halo = std::max(halo, std::abs(B.h_off[(size_t)k * kMaxDiags + j]));
        d.H = stage_halo(halo, es);
        d.C = stage_chunk(halo, es);
        for (int j = 0; j < 5; ++j) {
            d.off[j] = j < nd ? B.h_off[(size_t)k * kMaxDiags + j] : 0;
            d.uv[j] = j < nd ? (T)B.h_uv[(size_t)k * kMaxDiags + j] : T(0);
        }
        d.cd = (T)B.h_cd[k];
        d.cdok = B.h_cdok[k];
        const int need = stage_need(halo, es);
        d.vec = B.h_kind[k] == SK_STENCIL && nd <= 5 && B.h_cdok[k] && need <= cap
                    ? ((long long)need * es > kStageMinDoubles * 8LL ? 2 : 1)
                    : B.h_kind[k] == SK_CSR && csr_fits[k] ? 3 : 0;
    }
    out = DevBuf(B.ctx, tv.size() * sizeof(TileVec<T>));
    upload_on(B.ctx->stream, out.p, tv.data(), tv.size() * sizeof(TileVec<T>));
    (es == 4 ? B.tva32 : B.tva64) = DevBuf(B.ctx, tv.size() * sizeof(TileVec<T>));
    if constexpr (sizeof(T) == 4) B.h_tv32 = std::move(tv);
    else B.h_tv64 = std::move(tv);
}

static void build_tiles(Batch& B, int tile_rows, int tile_rows_dia) {
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
            if (B.h_kind[k] != SK_CSR) {
                j = std::min(e, i + tile_rows_dia);
            } else if (B.h_rp[i + 1] - base > kLdsNnz) {
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
        upload_on(B.ctx->stream, d.p, h.data(), h.size() * sizeof(int));
    };
    up(B.tile_r0, r0);
    up(B.tile_r1, r1);
    up(B.tile_sys, sys);
    build_tile_records<float>(B, r0, r1, sys, B.tv32);
    build_tile_records<double>(B, r0, r1, sys, B.tv64);
    static_assert(sizeof(TileVec<float>) == sizeof(TileVec<double>), "one staging size for both precisions");
    const size_t pb = sizeof(TileVec<double>) * (size_t)std::max(B.T, 1);
    if (pb > B.pin_act_bytes) {
        for (void*& q : B.pin_act) {
            if (q) HSPS_HIP(hipHostFree(q));
            q = nullptr;
        }
        B.pin_act_bytes = 0;
        for (void*& q : B.pin_act) HSPS_HIP(hipHostMalloc(&q, pb, hipHostMallocDefault));
        B.pin_act_bytes = pb;
    }
    up(B.sys_t0, B.h_sys_t0);
    up(B.sys_t1, B.h_sys_t1);
    B.partials = DevBuf(B.ctx, sizeof(double) * 2 * (size_t)std::max(B.T, 1));
    B.tagp = DevBuf(B.ctx, sizeof(unsigned long long) * 8 * (size_t)std::max(B.T, 1));
    B.tile_rows = tile_rows;
    B.tile_rows_dia = tile_rows_dia;
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

// Storage of one system with rows [a, e) of the global CSR (columns global):
// its kind and tables, the presence bytes of its rows (mask, e - a entries)
// and, for SK_DIA, its diagonal-major values (nd * (e - a), zero where absent).
struct SysStorage {
    int kind = SK_CSR, nd = 0, cdok = 0;
    int off[kMaxDiags] = {0, 0, 0, 0, 0, 0, 0, 0};
    double uv[kMaxDiags] = {0, 0, 0, 0, 0, 0, 0, 0};
    double cd = 0.0;
    std::vector<unsigned char> mask;
    std::vector<double> dia;
};

static int storage_cap() {
    // HSPS_BATCH_STORAGE=csr keeps every system in CSR; =dia allows DIA but not stencil
    const char* cap_env = std::getenv("HSPS_BATCH_STORAGE");
    return !cap_env ? SK_STENCIL : !std::strcmp(cap_env, "csr") ? SK_CSR : !std::strcmp(cap_env, "dia") ? SK_DIA : SK_STENCIL;
}

static SysStorage analyze_system(int a, int e, const int* rp, const int* ci, const double* v, const double* dinv,
                                 int cap) {
    SysStorage st;
    st.mask.assign((size_t)(e - a), 0);
    if (e > a) {
        bool same = true;
        for (int i = a + 1; i < e && same; ++i) same = dinv[i] == dinv[a];
        if (same) {
            st.cdok = 1;
            st.cd = dinv[a];
        }
    }
    const long long nz = rp[e] - rp[a];
    std::vector<int> o;
    bool ok = e > a;
    for (int i = a; i < e && ok; ++i)
        for (int q = rp[i]; q < rp[i + 1] && ok; ++q) {
            const int d = ci[q] - i;
            if (std::find(o.begin(), o.end(), d) == o.end()) {
                if ((int)o.size() == kMaxDiags) ok = false;
                else o.push_back(d);
            }
        }
    if (!ok || cap == SK_CSR) return st;
    std::sort(o.begin(), o.end());
    const int m = (int)o.size();
    st.nd = m;
    for (int d = 0; d < m; ++d) st.off[d] = o[d];
    // one value per diagonal, no repeated column in a row?
    std::vector<char> seen(m, 0);
    bool uniform = true;
    for (int i = a; i < e && uniform; ++i) {
        unsigned bits = 0;
        for (int q = rp[i]; q < rp[i + 1] && uniform; ++q) {
            const int d = (int)(std::lower_bound(o.begin(), o.end(), ci[q] - i) - o.begin());
            if (bits & (1u << d)) uniform = false;
            bits |= 1u << d;
            double& u = st.uv[d];
            if (!seen[d]) { seen[d] = 1; u = v[q]; }
            else if (!(v[q] == u) || std::signbit(v[q]) != std::signbit(u)) uniform = false;
        }
        st.mask[i - a] = (unsigned char)bits;
    }
    if (uniform && cap == SK_STENCIL && st.cdok) {  // (one diagonal value: D^-1 is a scalar)
        st.kind = SK_STENCIL;
        return st;
    }
    std::fill(st.mask.begin(), st.mask.end(), (unsigned char)0);
    for (int d = 0; d < kMaxDiags; ++d) st.uv[d] = 0.0;
    if ((long long)m * (e - a) * 2 > nz * 3) {
        st.nd = 0;
        return st;
    }
    st.kind = SK_DIA;
    const int n_s = e - a;
    st.dia.assign((size_t)m * n_s, 0.0);
    for (int i = a; i < e; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int d = (int)(std::lower_bound(o.begin(), o.end(), ci[q] - i) - o.begin());
            st.dia[(size_t)d * n_s + (i - a)] += v[q];
        }
    return st;
}

// The LDS staging window (doubles) a diagonal-stored system's chunk and halo
// take in the precision with element size es; 0 when no window up to
// kStageMaxDoubles holds them (rows then come from memory).
static int stage_doubles_of(const SysStorage& st, int es) {
    if (st.kind == SK_CSR) return 0;
    int h = 0;
    for (int d = 0; d < st.nd; ++d) h = std::max(h, std::abs(st.off[d]));
    const long long dbl = ((long long)stage_need(h, es) * es + 7) / 8;
    return dbl <= kStageMaxDoubles ? (int)dbl : 0;
}

// Storage selection per system.  A system whose nonzeros lie on <= kMaxDiags
// diagonals is diagonal-stored: SK_STENCIL when every diagonal carries a
// single value and no row repeats a column (constant-coefficient stencils:
// one value per diagonal plus a presence bit per row), else SK_DIA when the
// padding stays under 50% (diagonal-major values, zero where absent); all
// others stay CSR.  The iteration kernels of diagonal-stored systems read no
// column indices or row pointers.  Independently, a system whose Jacobi
// diagonal D^-1 is one value everywhere gets it as a per-system scalar.
static void build_dia(Batch& B, const int* row_off, const int* rp, const int* ci, const double* v,
                      const std::vector<double>& dinv) {
    const int S = B.S;
    std::vector<int> kind(S, SK_CSR), nd(S, 0), offs((size_t)S * kMaxDiags, 0), cdok(S, 0);
    std::vector<long long> base(S, 0);
    std::vector<double> uv((size_t)S * kMaxDiags, 0.0), cd(S, 0.0);
    std::vector<unsigned char> mask((size_t)B.N + kPad, 0);
    const int cap = storage_cap();
    std::vector<SysStorage> sto(S);
    long long total = 0;
    for (int k = 0; k < S; ++k) {
        const int s = B.order[k];
        const int a = row_off[s], e = row_off[s + 1];
        sto[k] = analyze_system(a, e, rp, ci, v, dinv.data(), cap);
        const SysStorage& st = sto[k];
        kind[k] = st.kind;
        nd[k] = st.nd;
        cdok[k] = st.cdok;
        cd[k] = st.cd;
        for (int d = 0; d < kMaxDiags; ++d) {
            offs[(size_t)k * kMaxDiags + d] = st.off[d];
            uv[(size_t)k * kMaxDiags + d] = st.uv[d];
        }
        std::copy(st.mask.begin(), st.mask.end(), mask.begin() + a);
        if (st.kind == SK_DIA) {
            base[k] = total;
            total += (long long)st.dia.size();
        }
    }
    // LDS window per precision: the largest halo of a tiled diagonal-stored system
    for (int prec = 0; prec < 2; ++prec) {
        const int es = prec ? 4 : 8;
        int need = kStageMinDoubles;
        for (int k = 0; k < std::min(B.n_large, S); ++k) need = std::max(need, stage_doubles_of(sto[k], es));
        B.stage_doubles[prec] = need;
    }
    std::vector<double> d64((size_t)std::max<long long>(total, 1), 0.0);
    for (int k = 0; k < S; ++k) {
        ++B.n_kind[kind[k]];
        if (kind[k] == SK_DIA) std::copy(sto[k].dia.begin(), sto[k].dia.end(), d64.begin() + base[k]);
    }
    auto up = [&](DevBuf& dst, const void* src, size_t bytes) {
        dst = DevBuf(B.ctx, std::max<size_t>(bytes, 8));
        upload_on(B.ctx->stream, dst.p, src, bytes);
    };
    B.h_kind = kind;
    B.h_nd = nd;
    B.h_off = offs;
    B.h_cdok = cdok;
    B.h_uv = uv;
    B.h_cd = cd;
    up(B.dia_kind, kind.data(), kind.size() * sizeof(int));
    up(B.dia_nd, nd.data(), nd.size() * sizeof(int));
    up(B.dia_off, offs.data(), offs.size() * sizeof(int));
    up(B.dia_base, base.data(), base.size() * sizeof(long long));
    up(B.dia64, d64.data(), (size_t)total * sizeof(double));
    std::vector<float> d32(d64.begin(), d64.begin() + total);
    up(B.dia32, d32.data(), (size_t)total * sizeof(float));
    up(B.uv64, uv.data(), uv.size() * sizeof(double));
    std::vector<float> u32(uv.begin(), uv.end());
    up(B.uv32, u32.data(), u32.size() * sizeof(float));
    up(B.umask, mask.data(), mask.size());
    up(B.cd_ok, cdok.data(), cdok.size() * sizeof(int));
    up(B.cd64, cd.data(), cd.size() * sizeof(double));
    std::vector<float> c32(cd.begin(), cd.end());
    up(B.cd32, c32.data(), c32.size() * sizeof(float));
}

static BArgs args_of(const Batch& B, const hsps_params& p) {
    BArgs a{};  // value-initialised: padding is zero, so argument blocks compare bytewise
    a.tv32 = B.tv32.as<TileVec<float>>();
    a.tv64 = B.tv64.as<TileVec<double>>();
    a.tva32 = B.tva32.as<TileVec<float>>();
    a.tva64 = B.tva64.as<TileVec<double>>();
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
    a.dia_kind = B.dia_kind.as<int>();
    a.dia_nd = B.dia_nd.as<int>();
    a.dia_off = B.dia_off.as<int>();
    a.dia_base = B.dia_base.as<long long>();
    a.dia64 = B.dia64.as<double>();
    a.dia32 = B.dia32.as<float>();
    a.uv64 = B.uv64.as<double>();
    a.uv32 = B.uv32.as<float>();
    a.umask = B.umask.as<unsigned char>();
    a.cd_ok = B.cd_ok.as<int>();
    a.cd64 = B.cd64.as<double>();
    a.cd32 = B.cd32.as<float>();
    a.stage_doubles = B.stage_doubles[p.precision == HSPS_MIXED ? 1 : 0];
    a.st = B.st.as<BState>();
    a.partials = B.partials.as<double>();
    a.tagp = B.tagp.as<unsigned long long>();
    a.T = B.T;
    a.tol = p.tol;
    a.delta = p.reliable_delta > 0 ? p.reliable_delta : 0.1;
    a.max_iter = p.max_iter;
    return a;
}

// The active set of a wave: the tiles K1 / K2 run (their records, compacted)
// and the number of their systems.  kinds: bit 0 / bit 2 = some tile takes the
// narrow / wide five-point kernels, bit 3 the CSR ones, bit 1 the general ones.
struct Active {
    int nt = 0, ns = 0, kinds = 0;
};
// bit of Active::kinds for a tile record's kernel family: 1 narrow and 4 wide
// five-point, 8 CSR (vec = 3), 2 the general kernels
static int kind_bit(int vec) { return vec == 1 ? 1 : vec == 2 ? 4 : vec == 3 ? 8 : 2; }

template <class T>
static const std::vector<TileVec<T>>& host_records(const Batch& B) {
    if constexpr (sizeof(T) == 4) return B.h_tv32;
    else return B.h_tv64;
}

// The dispatcher starts a launch's workgroups in block order and block b
// takes active record xcd_tile(b): mark, per system, the record of its
// last-started block (the one that runs the fused finish).
static int xcd_tile_host(int b, int nt) {
    constexpr int span = kXcds * kXcdRun;
    if (b >= nt / span * span) return b;
    const int x = b % kXcds, slot = b / kXcds;
    return ((slot / kXcdRun) * kXcds + x) * kXcdRun + slot % kXcdRun;
}
template <class T>
static void mark_finishers(TileVec<T>* rec, int nt, int n_sys) {
    std::vector<int> last((size_t)n_sys, -1);  // system -> record of its last-started block
    for (int b = 0; b < nt; ++b) {
        const int p = xcd_tile_host(b, nt);
        rec[p].fin = 0;
        last[rec[p].sys] = p;
    }
    for (int p : last)
        if (p >= 0) rec[p].fin = 1;
}

// Stages the records of the tiles of the live systems of [k0, k1) (live =
// nullptr: all of them), marks each system's finisher, and copies them to the
// device on stream s behind the work already queued there.
template <class T>
static Active upload_active(Batch& B, int k0, int k1, const std::vector<char>* live, hipStream_t s) {
    const std::vector<TileVec<T>>& tv = host_records<T>(B);
    char* base = (char*)B.pin_act[B.pin_next];
    B.pin_next ^= 1;
    TileVec<T>* rec = (TileVec<T>*)base;
    Active A;
    for (int k = k0; k < k1; ++k) {
        if (live && !(*live)[k - k0]) continue;
        ++A.ns;
        for (int t = B.h_sys_t0[k]; t < B.h_sys_t1[k]; ++t) {
            rec[A.nt++] = tv[t];
            A.kinds |= kind_bit(tv[t].vec);
        }
    }
    mark_finishers(rec, A.nt, B.S);
    if (A.nt > 0) {
        HSPS_HIP(hipMemcpyAsync((sizeof(T) == 4 ? B.tva32 : B.tva64).p, rec, sizeof(TileVec<T>) * (size_t)A.nt,
                                hipMemcpyHostToDevice, s));
    }
    return A;
}

template <class T, bool MIXED>
static void enqueue_iters(const Batch& B, const BArgs& a, const Active& act, int iters, const double* b, double* x,
                          hipStream_t s) {
    T* r = B.r.as<T>();
    T* p0 = B.p.as<T>();
    T* p1 = B.p1.as<T>();
    T* q = B.q.as<T>();
    float* xl = B.xlo.as<float>();
    const int nt = act.nt;
    const bool narrow = act.kinds & 1, any_gen = act.kinds & 2, wide = act.kinds & 4, csr = act.kinds & 8;
    for (int it = 0; it < iters; ++it) {
        const size_t lds = sizeof(double) * (size_t)a.stage_doubles;
        if (narrow)
            hipLaunchKernelGGL((k_b_k1v<T, MIXED, kVecNarrow>), dim3(nt), dim3(kVecNarrow), lds, s, a, p0, p1, r, b,
                               (const double*)x, (const float*)xl);
        if (wide)
            hipLaunchKernelGGL((k_b_k1v<T, MIXED, kVecWide>), dim3(nt), dim3(kVecWide), lds, s, a, p0, p1, r, b,
                               (const double*)x, (const float*)xl);
        if (any_gen)
            hipLaunchKernelGGL((k_b_k1<T, MIXED>), dim3(nt), dim3(kBlock), lds, s, a, p0, p1, q, r, b,
                               (const double*)x, (const float*)xl);
        if (csr)
            hipLaunchKernelGGL((k_b_k1c<T, MIXED>), dim3(nt), dim3(kBlock), lds, s, a, p0, p1, q, r, b,
                               (const double*)x, (const float*)xl);
        if (narrow)
            hipLaunchKernelGGL((k_b_k2v<T, MIXED, kVecNarrow>), dim3(nt), dim3(kVecNarrow), lds, s, a, (const T*)p0,
                               (const T*)p1, r, x, xl);
        if (wide)
            hipLaunchKernelGGL((k_b_k2v<T, MIXED, kVecWide>), dim3(nt), dim3(kVecWide), lds, s, a, (const T*)p0,
                               (const T*)p1, r, x, xl);
        if (any_gen)
            hipLaunchKernelGGL((k_b_k2<T, MIXED>), dim3(nt), dim3(kBlock), lds, s, a, (const T*)p0, (const T*)p1,
                               (const T*)q, r, x, xl);
        if (csr)
            hipLaunchKernelGGL((k_b_k2c<T, MIXED>), dim3(nt), dim3(kBlock), 0, s, a, (const T*)p0, (const T*)p1,
                               (const T*)q, r, x, xl);
    }
    HSPS_CHECK_LAUNCH();
}

static void ensure_work(Batch& B, int prec) {
    const size_t n = (size_t)B.N + kPad;  // 16-byte groups may run past the last row
    if (!B.st.p) {
        B.st = DevBuf(B.ctx, sizeof(BState) * (size_t)std::max(B.S, 1));
        B.r64 = DevBuf(B.ctx, n * sizeof(double));
    }
    if (B.work_prec != prec) {
        const size_t vb = prec == HSPS_MIXED ? 4 : 8;
        B.r = DevBuf(B.ctx, n * vb);
        B.p = DevBuf(B.ctx, n * vb);
        B.p1 = DevBuf(B.ctx, n * vb);
        B.q = DevBuf(B.ctx, n * vb);
        B.xlo = DevBuf(B.ctx, prec == HSPS_MIXED ? n * sizeof(float) : 0);
        B.work_prec = prec;
        B.drop_graphs();
    }
}

constexpr size_t kMaxGraphs = 16;  // captured chunks kept per batch

static void launch_chunk(Batch& B, const BArgs& a, int w, int chunk, const Active& act, const double* b, double* x,
                         int prec, hipStream_t s, bool use_graph) {
    if (act.nt <= 0) return;
    auto direct = [&]() {
        if (prec == HSPS_MIXED) enqueue_iters<float, true>(B, a, act, chunk, b, x, s);
        else enqueue_iters<double, false>(B, a, act, chunk, b, x, s);
    };
    if (!use_graph) { direct(); return; }
    Batch::G* g = nullptr;
    for (auto& c : B.graphs)
        if (c.wave == w && c.prec == prec && c.chunk == chunk && c.nt == act.nt && c.ns == act.ns &&
            c.kinds == act.kinds && std::memcmp(&c.args, &a, sizeof(BArgs)) == 0)
            g = &c;
    if (!g) {
        if (B.graphs.size() >= kMaxGraphs) {  // oldest first
            (void)hipGraphExecDestroy(B.graphs.front().exec);
            B.graphs.erase(B.graphs.begin());
        }
        // A capture that fails (another thread of the process may invalidate
        // it, e.g. by synchronising the device) records nothing that runs:
        // the chunk is then launched directly, and the failure counted.
        hipGraph_t graph = nullptr;
        HSPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        bool captured = true;
        try {
            direct();
        } catch (const Error&) {
            captured = false;
        }
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        Batch::G ng;
        hipError_t ei = hipErrorUnknown;
        if (captured && ec == hipSuccess && graph) ei = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            // an invalidated capture may leave the stream capturing: end it
            // before launching directly
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(s, &g2);
                if (g2) (void)hipGraphDestroy(g2);
            }
            (void)hipGetLastError();
            ++B.capture_failures;
            direct();
            return;
        }
        ng.wave = w;
        ng.prec = prec;
        ng.chunk = chunk;
        ng.nt = act.nt;
        ng.ns = act.ns;
        ng.kinds = act.kinds;
        ng.args = a;
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
        hipLaunchKernelGGL((k_b_finish<2, OpInit1>), dim3(B.n_large), dim3(kFinishThreads), 0, s, a, 0, (const int*)nullptr, OpInit1{});
        if (prec == HSPS_MIXED)
            hipLaunchKernelGGL((k_b_init2<float, true>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                               (const double*)B.r64.as<double>(), B.r.as<float>(), B.p.as<float>(), B.xlo.as<float>());
        else
            hipLaunchKernelGGL((k_b_init2<double, false>), dim3(B.T), dim3(kBlock), 0, s, a, 0,
                               (const double*)B.r64.as<double>(), B.r.as<double>(), B.p.as<double>(), (float*)nullptr);
        hipLaunchKernelGGL((k_b_finish<1, OpInit2>), dim3(B.n_large), dim3(kFinishThreads), 0, s, a, 0, (const int*)nullptr, OpInit2{});
        HSPS_CHECK_LAUNCH();
    }
    const int chunk = std::max(1, std::min(p.check_every > 0 ? p.check_every : 32, 64));
    if (use_graph && (b != B.graph_b || x != B.graph_x)) {  // captured kernels hold b and x
        B.drop_graphs();
        B.graph_b = b;
        B.graph_x = x;
    }
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
            // the kernels run the tiles and systems still iterating: every one at
            // first, and the halt flags the host reads (one chunk behind) shrink
            // the set whenever its live tiles are at most half of those launched
            auto upload = [&](const std::vector<char>* live) {
                return prec == HSPS_MIXED ? upload_active<float>(B, k0, k1, live, s)
                                          : upload_active<double>(B, k0, k1, live, s);
            };
            Active act = upload(nullptr);
            std::vector<char> live(ns, 1);
            for (;;) {
                launch_chunk(B, a, w, chunk, act, b, x, prec, s, use_graph);
                ++chunks;
                BState* slot = pin + (size_t)(k & 1) * B.S;
                HSPS_HIP(hipMemcpyAsync(slot + k0, B.st.as<BState>() + k0, sizeof(BState) * ns, hipMemcpyDeviceToHost, s));
                HSPS_HIP(hipEventRecord(ev[k & 1], s));
                bool all = false;
                if (pending) {
                    HSPS_HIP(hipEventSynchronize(ev[(k - 1) & 1]));
                    const BState* prev = pin + (size_t)((k - 1) & 1) * B.S;
                    all = true;
                    int live_tiles = 0;
                    for (int i = k0; i < k1; ++i) {
                        live[i - k0] = prev[i].halt == 0;
                        if (live[i - k0]) {
                            all = false;
                            live_tiles += B.h_sys_t1[i] - B.h_sys_t0[i];
                        }
                    }
                    if (!all && 2 * live_tiles <= act.nt) act = upload(&live);
                }
                pending = true;
                ++k;
                if (all || chunks >= cap || (max_chunks > 0 && chunks >= max_chunks)) break;
            }
        }
        if (B.T > 0) {
            if (prec == HSPS_MIXED)  // every large system's rows, one workgroup per tile
                hipLaunchKernelGGL(k_b_flush_tiles, dim3(B.T), dim3(kBlock), 0, s, a, 0, x, B.xlo.as<float>());
            hipLaunchKernelGGL(k_b_final, dim3(B.T), dim3(kBlock), 0, s, a, 0, b, (const double*)x);
            hipLaunchKernelGGL((k_b_finish<1, OpFinal>), dim3(B.n_large), dim3(kFinishThreads), 0, s, a, 0, (const int*)nullptr, OpFinal{});
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

// Tile-size autotuner: times a few chunks of the tiled waves for each
// candidate on a scratch copy of x and keeps the fastest — first the tile of
// diagonal-stored systems, then the CSR tile, each only when the batch has
// large systems of that kind.
static void autotune(Batch& B, const double* b, const double* x0, const hsps_params& p) {
    bool has_csr = false, has_dia = false;
    for (int k = 0; k < B.n_large; ++k) (B.h_kind[k] == SK_CSR ? has_csr : has_dia) = true;
    int best_c = 512, best_d = 4096;
    if (B.n_large == 0) {
        build_tiles(B, best_c, best_d);
        return;
    }
    hipStream_t s = B.ctx->stream;
    DevBuf xs(B.ctx, (size_t)std::max(B.N, 1) * sizeof(double));
    hsps_params q = p;
    q.check_every = 16;
    q.max_iter = 48;  // bounds the small-system kernel during tuning
    auto trial = [&](int c, int d) {
        build_tiles(B, c, d);
        HSPS_HIP(hipMemcpyAsync(xs.p, x0, (size_t)B.N * sizeof(double), hipMemcpyDeviceToDevice, s));
        (void)run_batch(B, b, xs.as<double>(), q, false, 1);  // warm
        HSPS_HIP(hipMemcpyAsync(xs.p, x0, (size_t)B.N * sizeof(double), hipMemcpyDeviceToDevice, s));
        HSPS_HIP(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        (void)run_batch(B, b, xs.as<double>(), q, false, 2);
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    if (has_dia) {
        double bt = 1e30;
        for (int d : {1024, 2048, 4096, 8192}) {
            const double t = trial(best_c, d);
            if (t < bt) { bt = t; best_d = d; }
        }
    }
    if (has_csr) {
        double bt = 1e30;
        for (int c : {256, 512, 1024}) {
            const double t = trial(c, best_d);
            if (t < bt) { bt = t; best_c = c; }
        }
    }
    build_tiles(B, best_c, best_d);
}

// Validates the concatenated systems and builds B on ctx: storage per
// system, tiles (CSR 256 rows, diagonal-stored 1024; retuned on the first
// solve), waves.
static void fill_batch(Batch& B, Ctx* ctx, int S, const int* row_off, const int* rp, const int* ci, const double* v) {
    if (!ctx || S < 0 || !row_off || !rp) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: bad argument");
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
    B.ctx = ctx;
    B.S = S;
    B.N = N;
    B.nnz = nnz;
    B.h_row_off.assign(row_off, row_off + S + 1);
    B.h_rp.assign(rp, rp + N + 1);
    ctx->set_current();
    HSPS_HIP(hipStreamCreateWithFlags(&B.aux, hipStreamNonBlocking));
    HSPS_HIP(hipEventCreateWithFlags(&B.join, hipEventDisableTiming));
    HSPS_HIP(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
    B.rp = DevBuf(ctx, sizeof(int) * (size_t)(N + 1));
    B.ci = DevBuf(ctx, sizeof(int) * (size_t)std::max<long long>(nnz, 1));
    B.v64 = DevBuf(ctx, sizeof(double) * (size_t)std::max<long long>(nnz, 1));
    B.v32 = DevBuf(ctx, sizeof(float) * (size_t)std::max<long long>(nnz, 1));
    B.dinv64 = DevBuf(ctx, sizeof(double) * (size_t)std::max(N, 1));
    B.dinv32 = DevBuf(ctx, sizeof(float) * (size_t)std::max(N, 1));
    upload_on(ctx->stream, B.rp.p, rp, sizeof(int) * (size_t)(N + 1));
    if (nnz) {
        upload_on(ctx->stream, B.ci.p, ci, sizeof(int) * (size_t)nnz);
        upload_on(ctx->stream, B.v64.p, v, sizeof(double) * (size_t)nnz);
        std::vector<float> f(v, v + nnz);
        upload_on(ctx->stream, B.v32.p, f.data(), sizeof(float) * (size_t)nnz);
    }
    if (N) {
        upload_on(ctx->stream, B.dinv64.p, d64.data(), sizeof(double) * (size_t)N);
        std::vector<float> f(d64.begin(), d64.end());
        upload_on(ctx->stream, B.dinv32.p, f.data(), sizeof(float) * (size_t)N);
    }
    B.order.resize(S);
    std::iota(B.order.begin(), B.order.end(), 0);
    std::stable_sort(B.order.begin(), B.order.end(), [&](int a, int b) {
        return rp[row_off[a + 1]] - rp[row_off[a]] > rp[row_off[b + 1]] - rp[row_off[b]];
    });
    B.n_large = 0;
    for (int k = 0; k < S; ++k)
        if (row_off[B.order[k] + 1] - row_off[B.order[k]] > kSmallRows) B.n_large = k + 1;
    std::vector<int> r0(S), r1(S);
    for (int k = 0; k < S; ++k) {
        r0[k] = row_off[B.order[k]];
        r1[k] = row_off[B.order[k] + 1];
    }
    B.sys_row0 = DevBuf(ctx, sizeof(int) * (size_t)std::max(S, 1));
    B.sys_row1 = DevBuf(ctx, sizeof(int) * (size_t)std::max(S, 1));
    if (S) {
        upload_on(ctx->stream, B.sys_row0.p, r0.data(), sizeof(int) * S);
        upload_on(ctx->stream, B.sys_row1.p, r1.data(), sizeof(int) * S);
    }
    build_dia(B, row_off, rp, ci, v, d64);
    build_tiles(B, 256, 1024);
    B.tile_rows = 0;  // not tuned yet
    // Wave working-set budget (MB).  Measured on MI355X (512 systems,
    // 1.9 TB of CG traffic): one wave beats MALL-sized (<=192 MB)
    // waves — the fixed per-iteration launch cost of extra waves
    // outweighs the cache hits — so the default is a single wave.
    const char* wmb = std::getenv("HSPS_BATCH_WAVE_MB");
    const size_t budget_mb = wmb ? (size_t)std::max(8, std::atoi(wmb)) : ((size_t)1 << 20);
    build_waves(B, budget_mb << 20);
}

bool solo_cg(Csr& A, const double* b, double* x, const hsps_params& p, SoloResult* out) {
    if (A.solo_use == 0) return false;
    if (!A.solo) {
        auto B = std::make_shared<Batch>();
        const int row_off[2] = {0, A.rows};
        fill_batch(*B, A.ctx, 1, row_off, A.h_rp.data(), A.h_ci.data(), A.h_v.data());
        // a large CSR-stored system: the plain kernel sequence (row-block SpMV) serves it better
        if (B->n_large > 0 && B->h_kind[0] == SK_CSR) {
            A.solo_use = 0;
            return false;
        }
        // fixed tiles (no autotuning trials inside a single solve): 2048 rows,
        // or one staging chunk of the system in both precisions when that is
        // longer (a wide halo: a 2048-row tile of a 2048-wide grid stages three
        // times its rows; a 4092-row tile twice; measured on MI355X, Poisson
        // 2048^2 mixed PCG: 83 -> 74 ms)
        int halo = 0;
        for (int d = 0; d < B->h_nd[0]; ++d) halo = std::max(halo, std::abs(B->h_off[d]));
        const int tr = std::max(2048, std::min(stage_chunk(halo, 4), stage_chunk(halo, 8)));
        build_tiles(*B, 512, tr);
        A.solo = B;
        A.solo_use = 1;
    }
    static const bool no_graph = std::getenv("HSPS_BATCH_NO_GRAPH") != nullptr;
    const std::vector<BState> st = run_batch(*A.solo, b, x, p, !no_graph && p.batch_graphs != 0);
    const BState& S = st[0];
    out->iterations = S.iter;
    out->converged = S.conv;
    out->restarts = S.restarts;
    out->restarts_pq = S.restarts_pq;
    out->reliable_updates = S.reliable;
    out->initial_residual = S.r0;
    out->final_residual = S.rel;
    out->true_residual = S.true_rel;
    return true;
}


// =============================================================================
// Sessions: systems join a running batch and leave it at chunk boundaries
// (no reference counterpart; the serving loop's continuous batching).
// =============================================================================
// A session holds the engine's device tables at a fixed capacity (rows,
// nonzeros, systems, tiles, DIA values).  A joining system gets its own
// ranges from free lists: rows [row0, row0 + n) and at least one gap row
// after them (its n + 1 row-pointer entries are its own), row0 a multiple of
// kSessRowAlign (the five-point kernels deal rows to threads by 16-byte
// groups, so a system's reductions keep their order only at a fixed
// alignment; a batch places its first system at row 0), a nonzero range, a system
// slot, a tile range and, if DIA-stored, a value range.  Its storage
// (analyze_system), tiles and tile records are the ones it would have alone,
// and every per-system reduction runs over its own tiles in a fixed order,
// so its iterates and result are the same bits whichever systems share the
// session and whenever it joins.  Every system is tiled here (the
// one-workgroup path of hsps_batch_solve would serialise joins on one
// stream); small systems take one tile.  A join uploads the system and
// enqueues its initialisation (r = b - A x0, rho) behind the chunk in
// flight; the next chunk's active list includes it.  The host reads the
// systems' states one chunk behind, as run_batch does.

constexpr int kSessRowAlign = 64;

class FreeList {
public:
    void reset(long long cap) { free_.assign(1, {0, cap}); }
    // first fit; -1 when no range of len is free (len 0 takes nothing)
    long long take(long long len) {
        if (len <= 0) return 0;
        for (size_t i = 0; i < free_.size(); ++i)
            if (free_[i].second >= len) {
                const long long s = free_[i].first;
                free_[i].first += len;
                free_[i].second -= len;
                if (free_[i].second == 0) free_.erase(free_.begin() + (long)i);
                return s;
            }
        return -1;
    }
    bool fits(long long len) const {
        for (const auto& f : free_)
            if (f.second >= len) return true;
        return len <= 0;
    }
    void give(long long s, long long len) {
        if (len <= 0) return;
        size_t i = 0;
        while (i < free_.size() && free_[i].first < s) ++i;
        free_.insert(free_.begin() + (long)i, {s, len});
        if (i + 1 < free_.size() && free_[i].first + free_[i].second == free_[i + 1].first) {
            free_[i].second += free_[i + 1].second;
            free_.erase(free_.begin() + (long)i + 1);
        }
        if (i > 0 && free_[i - 1].first + free_[i - 1].second == free_[i].first) {
            free_[i - 1].second += free_[i].second;
            free_.erase(free_.begin() + (long)i);
        }
    }

private:
    std::vector<std::pair<long long, long long>> free_;  // (start, length), ascending, disjoint
};

struct SessSlot {
    int used = 0;      // holds a system (joined, not yet released)
    int live = 0;      // still iterating as far as the host knows
    int done = 0;      // its halt has been reported
    int conv = 0;      // converged, as read back at the halt
    long long join_seq = 0;  // the first state read-back that includes its initialisation
    int row0 = 0, n = 0, row_len = 0;  // rows [row0, row0 + row_len) held, n of them the system's
    long long nz0 = 0, nnz = 0, dia0 = 0, dia_len = 0;
    int t0 = 0, t1 = 0;
    int need[2] = {0, 0};  // LDS window in doubles, FP64 / FP32
};

struct Session {
    Batch B;  // device tables at capacity: B.S slots, B.N rows (gap rows included), B.T tiles
    FreeList rows, nzs, tiles, dias;
    std::vector<int> free_slots;
    std::vector<SessSlot> slots;
    hsps_params p{};
    int prec = HSPS_FP64;
    int tile_rows = 512, tile_rows_dia = 2048;
    long long nnz_cap = 0, dia_cap = 0;
    DevBuf b, x;  // right-hand sides and solutions by row (kPad rows of slack)
    bool dirty = true;
    Active act;
    int stage = kStageMinDoubles;
    BState* pin = nullptr;  // two read-back slots of slots.size() states
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long seq = 0;      // read-backs enqueued so far
    bool pending = false;
    std::vector<std::pair<std::array<int, 4>, int>> keys;  // chunk shapes seen (a graph on the second use)
    ~Session() {
        if (pin) (void)hipHostFree(pin);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Tile rows of diagonal-stored systems for sessions: the batch autotuner's
// choice on a probe batch of five-point systems (8 of 300^2 and 4 of 600^2
// rows, Poisson), made once per device and precision in a process, so every
// session of the process tiles a system the same way (same bits).
static int probe_tile_rows_dia(Ctx* ctx, const hsps_params& p) {
    std::vector<int> sides = {300, 300, 300, 300, 300, 300, 300, 300, 600, 600, 600, 600};
    std::vector<int> row_off(1, 0), rp(1, 0), ci;
    std::vector<double> v;
    for (int side : sides) {
        const int base = row_off.back();
        for (int i = 0; i < side; ++i)
            for (int j = 0; j < side; ++j) {
                const int row = i * side + j;
                auto push = [&](int c, double a) {
                    ci.push_back(base + c);
                    v.push_back(a);
                };
                if (i > 0) push(row - side, -1.0);
                if (j > 0) push(row - 1, -1.0);
                push(row, 4.0);
                if (j < side - 1) push(row + 1, -1.0);
                if (i < side - 1) push(row + side, -1.0);
                rp.push_back((int)ci.size());
            }
        row_off.push_back(base + side * side);
    }
    const int N = row_off.back();
    Batch B;
    fill_batch(B, ctx, (int)sides.size(), row_off.data(), rp.data(), ci.data(), v.data());
    ensure_work(B, p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64);
    DevBuf bd(ctx, sizeof(double) * (size_t)N), xd(ctx, sizeof(double) * (size_t)N);
    std::vector<double> ones((size_t)N, 1.0);
    upload_on(ctx->stream, bd.p, ones.data(), sizeof(double) * (size_t)N);
    fill_on(ctx->stream, xd.p, 0, sizeof(double) * (size_t)N);
    autotune(B, bd.as<double>(), xd.as<double>(), p);
    HSPS_HIP(hipStreamSynchronize(ctx->stream));
    return B.tile_rows_dia;
}

static int session_tile_rows_dia(Ctx* ctx, const hsps_params& p) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, int> chosen;  // (device, precision) -> tile rows
    const std::pair<int, int> key = {ctx->device, p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64};
    std::lock_guard<std::mutex> lk(mu);
    auto it = chosen.find(key);
    if (it != chosen.end()) return it->second;
    const int t = probe_tile_rows_dia(ctx, p);
    chosen[key] = t;
    return t;
}

static void session_alloc(Session& S, Ctx* ctx, int max_systems, long long max_rows, long long max_nnz,
                          const hsps_params& p) {
    if (max_systems < 1 || max_rows < 1 || max_nnz < 1) throw Error(HSPS_ERR_INVALID, "hsps_session_create: bad capacity");
    S.p = p;
    S.prec = p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64;
    if (p.batch_tile_rows > 0) {
        S.tile_rows = S.tile_rows_dia = std::min(p.batch_tile_rows, 1 << 16);
    } else {
        S.tile_rows = 512;
        S.tile_rows_dia = session_tile_rows_dia(ctx, p);
    }
    Batch& B = S.B;
    const long long N = max_rows + (long long)max_systems * kSessRowAlign;  // gap rows and alignment
    if (N + kPad >= (1LL << 31) || max_nnz >= (1LL << 31)) throw Error(HSPS_ERR_INVALID, "hsps_session_create: capacity beyond int32");
    // a CSR tile ends at tile_rows rows or when the next row would take it past
    // kLdsNnz nonzeros, so two consecutive tiles hold more than kLdsNnz of them
    const long long T = N / std::min(S.tile_rows, S.tile_rows_dia) + 2 * (max_nnz / kLdsNnz) + 2LL * max_systems + 1;
    B.ctx = ctx;
    B.S = max_systems;
    B.N = (int)N;
    B.nnz = max_nnz;
    B.T = (int)T;
    B.n_large = max_systems;
    S.nnz_cap = max_nnz;
    S.dia_cap = max_nnz + max_nnz / 2 + 1;  // a DIA-stored system keeps <= 1.5 x its nonzeros
    ctx->set_current();
    HSPS_HIP(hipStreamCreateWithFlags(&B.aux, hipStreamNonBlocking));
    HSPS_HIP(hipEventCreateWithFlags(&B.join, hipEventDisableTiming));
    HSPS_HIP(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
    const size_t n = (size_t)N + kPad, Sn = (size_t)max_systems, Tn = (size_t)T;
    B.rp = DevBuf(ctx, sizeof(int) * (n + 1));
    B.ci = DevBuf(ctx, sizeof(int) * (size_t)max_nnz);
    B.v64 = DevBuf(ctx, sizeof(double) * (size_t)max_nnz);
    B.v32 = DevBuf(ctx, sizeof(float) * (size_t)max_nnz);
    B.dinv64 = DevBuf(ctx, sizeof(double) * n);
    B.dinv32 = DevBuf(ctx, sizeof(float) * n);
    B.umask = DevBuf(ctx, n);
    B.dia64 = DevBuf(ctx, sizeof(double) * (size_t)S.dia_cap);
    B.dia32 = DevBuf(ctx, sizeof(float) * (size_t)S.dia_cap);
    B.dia_kind = DevBuf(ctx, sizeof(int) * Sn);
    B.dia_nd = DevBuf(ctx, sizeof(int) * Sn);
    B.dia_off = DevBuf(ctx, sizeof(int) * Sn * kMaxDiags);
    B.dia_base = DevBuf(ctx, sizeof(long long) * Sn);
    B.uv64 = DevBuf(ctx, sizeof(double) * Sn * kMaxDiags);
    B.uv32 = DevBuf(ctx, sizeof(float) * Sn * kMaxDiags);
    B.cd_ok = DevBuf(ctx, sizeof(int) * Sn);
    B.cd64 = DevBuf(ctx, sizeof(double) * Sn);
    B.cd32 = DevBuf(ctx, sizeof(float) * Sn);
    B.sys_row0 = DevBuf(ctx, sizeof(int) * Sn);
    B.sys_row1 = DevBuf(ctx, sizeof(int) * Sn);
    B.sys_t0 = DevBuf(ctx, sizeof(int) * Sn);
    B.sys_t1 = DevBuf(ctx, sizeof(int) * Sn);
    B.tile_r0 = DevBuf(ctx, sizeof(int) * Tn);
    B.tile_r1 = DevBuf(ctx, sizeof(int) * Tn);
    B.tile_sys = DevBuf(ctx, sizeof(int) * Tn);
    B.tv32 = DevBuf(ctx, sizeof(TileVec<float>) * Tn);
    B.tv64 = DevBuf(ctx, sizeof(TileVec<double>) * Tn);
    B.tva32 = DevBuf(ctx, sizeof(TileVec<float>) * Tn);
    B.tva64 = DevBuf(ctx, sizeof(TileVec<double>) * Tn);
    B.partials = DevBuf(ctx, sizeof(double) * 2 * Tn);
    B.tagp = DevBuf(ctx, sizeof(unsigned long long) * 8 * Tn);
    B.h_tv32.assign(Tn, TileVec<float>{});
    B.h_tv64.assign(Tn, TileVec<double>{});
    const size_t pb = sizeof(TileVec<double>) * Tn;
    for (void*& q : B.pin_act) HSPS_HIP(hipHostMalloc(&q, pb, hipHostMallocDefault));
    B.pin_act_bytes = pb;
    ensure_work(B, S.prec);
    S.b = DevBuf(ctx, sizeof(double) * n);
    S.x = DevBuf(ctx, sizeof(double) * n);
    fill_on(ctx->stream, S.b.p, 0, sizeof(double) * n);
    fill_on(ctx->stream, S.x.p, 0, sizeof(double) * n);
    fill_on(ctx->stream, B.umask.p, 0, n);
    HSPS_HIP(hipHostMalloc((void**)&S.pin, sizeof(BState) * Sn * 2, hipHostMallocDefault));
    HSPS_HIP(hipEventCreateWithFlags(&S.ev[0], hipEventDisableTiming));
    HSPS_HIP(hipEventCreateWithFlags(&S.ev[1], hipEventDisableTiming));
    S.rows.reset(N);
    S.nzs.reset(max_nnz);
    S.tiles.reset(T);
    S.dias.reset(S.dia_cap);
    S.slots.assign(Sn, SessSlot{});
    S.free_slots.clear();
    for (int k = max_systems - 1; k >= 0; --k) S.free_slots.push_back(k);
    B.graph_b = S.b.as<double>();
    B.graph_x = S.x.as<double>();
}

template <class T>
static TileVec<T> session_record(const SysStorage& st, int slot, int t, int r0, int r1, int row0, int n,
                                 bool csr_fits) {
    constexpr int es = (int)sizeof(T);
    TileVec<T> d;
    std::memset(&d, 0, sizeof d);
    d.r0 = r0;
    d.r1 = r1;
    d.sys = slot;
    d.tile = t;
    d.row0 = row0;
    d.last = row0 + n - 1;
    int halo = 0;
    for (int j = 0; j < st.nd; ++j) halo = std::max(halo, std::abs(st.off[j]));
    d.H = stage_halo(halo, es);
    d.C = stage_chunk(halo, es);
    for (int j = 0; j < 5; ++j) {
        d.off[j] = j < st.nd ? st.off[j] : 0;
        d.uv[j] = j < st.nd ? (T)st.uv[j] : T(0);
    }
    d.cd = (T)st.cd;
    d.cdok = st.cdok;
    // the five-point path when the system's own window fits the largest one a
    // launch provides (a batch of this system alone decides the same way)
    const int need = stage_need(halo, es), cap = kStageMaxDoubles * 8 / es;
    d.vec = st.kind == SK_STENCIL && st.nd <= 5 && st.cdok && need <= cap
                ? ((long long)need * es > kStageMinDoubles * 8LL ? 2 : 1)
                : st.kind == SK_CSR && csr_fits ? 3 : 0;
    return d;
}

// Uploads one n x n system (local CSR) with right-hand side h_b and initial
// guess d_x0 (device, or null for zero) and enqueues its initialisation.
// Returns the slot, or -1 when the session has no room for it.
static int session_add(Session& S, int n, const int* rp, const int* ci, const double* v, const double* h_b,
                       const double* d_x0) {
    Batch& B = S.B;
    if (n < 1 || !rp || !ci || !v || !h_b) throw Error(HSPS_ERR_INVALID, "hsps_session_add: bad argument");
    const long long nnz = rp[n];
    if (rp[0] != 0 || nnz < 1) throw Error(HSPS_ERR_INVALID, "hsps_session_add: bad row_ptr");
    // every range is a multiple of kSessRowAlign rows long, so every start is aligned
    const long long row_len = ((long long)n + kSessRowAlign) / kSessRowAlign * kSessRowAlign;
    // no room: said before the O(nnz) checks and analysis (a caller retries as systems end)
    if (S.free_slots.empty() || !S.rows.fits(row_len) || !S.nzs.fits(nnz)) return -1;
    std::vector<double> dinv(n);
    for (int i = 0; i < n; ++i) {
        if (rp[i + 1] < rp[i]) throw Error(HSPS_ERR_INVALID, "hsps_session_add: row_ptr not monotone");
        double d = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] < 0 || ci[k] >= n) throw Error(HSPS_ERR_INVALID, "hsps_session_add: column out of range");
            if (ci[k] == i) d = v[k];
        }
        if (std::abs(d) < kEps)  // jacobi.cpp:28-30
            throw Error(HSPS_ERR_ZERO_DIAG, "JacobiPreconditioner: zero diagonal at row " + std::to_string(i));
        dinv[i] = 1.0 / d;
    }
    const SysStorage st = analyze_system(0, n, rp, ci, v, dinv.data(), storage_cap());
    // tiles (build_tiles' rule, local rows)
    std::vector<int> tr0, tr1;
    for (int i = 0; i < n;) {
        int j = i;
        if (st.kind != SK_CSR) {
            j = std::min(n, i + S.tile_rows_dia);
        } else if (rp[i + 1] - rp[i] > kLdsNnz) {
            j = i + 1;
        } else {
            while (j < n && j - i < S.tile_rows && rp[j + 1] - rp[i] <= kLdsNnz) ++j;
        }
        tr0.push_back(i);
        tr1.push_back(j);
        i = j;
    }
    const int nt = (int)tr0.size();
    bool csr_fits = true;
    for (int t = 0; t < nt; ++t) csr_fits = csr_fits && rp[tr1[t]] - rp[tr0[t]] <= kLdsNnz;
    const long long row0 = S.rows.take(row_len);
    if (row0 < 0) return -1;
    const long long nz0 = S.nzs.take(nnz);
    if (nz0 < 0) { S.rows.give(row0, row_len); return -1; }
    const long long t0 = S.tiles.take(nt);
    if (t0 < 0) { S.rows.give(row0, row_len); S.nzs.give(nz0, nnz); return -1; }
    const long long dia_len = (long long)st.dia.size();
    const long long dia0 = st.kind == SK_DIA ? S.dias.take(dia_len) : 0;
    if (dia0 < 0) {
        S.rows.give(row0, row_len); S.nzs.give(nz0, nnz); S.tiles.give(t0, nt);
        return -1;
    }
    const int slot = S.free_slots.back();
    S.free_slots.pop_back();
    // the system's rows, nonzeros and tables, copied on the side stream while
    // a chunk may run (it reads none of these ranges); all sources stay alive
    // until the one wait below
    std::vector<int> grp(n + 1), gci(nnz);
    for (int i = 0; i <= n; ++i) grp[i] = (int)(nz0 + rp[i]);
    for (long long k = 0; k < nnz; ++k) gci[k] = (int)(row0 + ci[k]);
    std::vector<float> v32(v, v + nnz), d32(dinv.begin(), dinv.end());
    auto put = [&](const DevBuf& dst, size_t off_bytes, const void* src, size_t bytes) {
        if (bytes) HSPS_HIP(hipMemcpyAsync((char*)dst.p + off_bytes, src, bytes, hipMemcpyHostToDevice, B.aux));
    };
    put(B.rp, sizeof(int) * row0, grp.data(), sizeof(int) * (n + 1));
    put(B.ci, sizeof(int) * nz0, gci.data(), sizeof(int) * nnz);
    put(B.v64, sizeof(double) * nz0, v, sizeof(double) * nnz);
    put(B.v32, sizeof(float) * nz0, v32.data(), sizeof(float) * nnz);
    put(B.dinv64, sizeof(double) * row0, dinv.data(), sizeof(double) * n);
    put(B.dinv32, sizeof(float) * row0, d32.data(), sizeof(float) * n);
    put(B.umask, row0, st.mask.data(), (size_t)n);
    if (st.kind == SK_DIA) {
        std::vector<float> f(st.dia.begin(), st.dia.end());
        put(B.dia64, sizeof(double) * dia0, st.dia.data(), sizeof(double) * dia_len);
        put(B.dia32, sizeof(float) * dia0, f.data(), sizeof(float) * dia_len);
    }
    const int kind = st.kind, nd = st.nd, cdok = st.cdok, r0i = (int)row0, r1i = (int)(row0 + n), t0i = (int)t0,
              t1i = (int)(t0 + nt);
    float uv32[kMaxDiags], cd32 = (float)st.cd;
    for (int d = 0; d < kMaxDiags; ++d) uv32[d] = (float)st.uv[d];
    put(B.dia_kind, sizeof(int) * slot, &kind, sizeof(int));
    put(B.dia_nd, sizeof(int) * slot, &nd, sizeof(int));
    put(B.dia_off, sizeof(int) * slot * kMaxDiags, st.off, sizeof(int) * kMaxDiags);
    put(B.dia_base, sizeof(long long) * slot, &dia0, sizeof(long long));
    put(B.uv64, sizeof(double) * slot * kMaxDiags, st.uv, sizeof(double) * kMaxDiags);
    put(B.uv32, sizeof(float) * slot * kMaxDiags, uv32, sizeof(float) * kMaxDiags);
    put(B.cd_ok, sizeof(int) * slot, &cdok, sizeof(int));
    put(B.cd64, sizeof(double) * slot, &st.cd, sizeof(double));
    put(B.cd32, sizeof(float) * slot, &cd32, sizeof(float));
    put(B.sys_row0, sizeof(int) * slot, &r0i, sizeof(int));
    put(B.sys_row1, sizeof(int) * slot, &r1i, sizeof(int));
    put(B.sys_t0, sizeof(int) * slot, &t0i, sizeof(int));
    put(B.sys_t1, sizeof(int) * slot, &t1i, sizeof(int));
    std::vector<int> g0(nt), g1(nt), gs(nt, slot);
    std::vector<TileVec<float>> r32(nt);
    std::vector<TileVec<double>> r64(nt);
    for (int t = 0; t < nt; ++t) {
        g0[t] = r0i + tr0[t];
        g1[t] = r0i + tr1[t];
        r32[t] = session_record<float>(st, slot, t0i + t, g0[t], g1[t], r0i, n, csr_fits);
        r64[t] = session_record<double>(st, slot, t0i + t, g0[t], g1[t], r0i, n, csr_fits);
        B.h_tv32[t0i + t] = r32[t];
        B.h_tv64[t0i + t] = r64[t];
    }
    put(B.tile_r0, sizeof(int) * t0, g0.data(), sizeof(int) * nt);
    put(B.tile_r1, sizeof(int) * t0, g1.data(), sizeof(int) * nt);
    put(B.tile_sys, sizeof(int) * t0, gs.data(), sizeof(int) * nt);
    put(B.tv32, sizeof(TileVec<float>) * t0, r32.data(), sizeof(TileVec<float>) * nt);
    put(B.tv64, sizeof(TileVec<double>) * t0, r64.data(), sizeof(TileVec<double>) * nt);
    put(S.b, sizeof(double) * row0, h_b, sizeof(double) * n);
    HSPS_HIP(hipStreamSynchronize(B.aux));  // the copies are done before the initialisation is enqueued
    hipStream_t s = B.ctx->stream;
    double* xr = S.x.as<double>() + row0;
    if (d_x0) HSPS_HIP(hipMemcpyAsync(xr, d_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    else HSPS_HIP(hipMemsetAsync(xr, 0, sizeof(double) * n, s));
    // initialisation behind the chunk in flight (run_batch's init, this system's tiles only)
    const BArgs a = args_of(B, S.p);
    const double* bd = S.b.as<double>();
    hipLaunchKernelGGL(k_b_init1, dim3(nt), dim3(kBlock), 0, s, a, t0i, bd, (const double*)S.x.as<double>(),
                       B.r64.as<double>());
    hipLaunchKernelGGL((k_b_finish<2, OpInit1>), dim3(1), dim3(kFinishThreads), 0, s, a, slot, (const int*)nullptr,
                       OpInit1{});
    if (S.prec == HSPS_MIXED)
        hipLaunchKernelGGL((k_b_init2<float, true>), dim3(nt), dim3(kBlock), 0, s, a, t0i,
                           (const double*)B.r64.as<double>(), B.r.as<float>(), B.p.as<float>(), B.xlo.as<float>());
    else
        hipLaunchKernelGGL((k_b_init2<double, false>), dim3(nt), dim3(kBlock), 0, s, a, t0i,
                           (const double*)B.r64.as<double>(), B.r.as<double>(), B.p.as<double>(), (float*)nullptr);
    hipLaunchKernelGGL((k_b_finish<1, OpInit2>), dim3(1), dim3(kFinishThreads), 0, s, a, slot, (const int*)nullptr,
                       OpInit2{});
    HSPS_CHECK_LAUNCH();
    SessSlot& sl = S.slots[slot];
    sl = SessSlot{};
    sl.used = sl.live = 1;
    sl.join_seq = S.seq;
    sl.row0 = r0i;
    sl.n = n;
    sl.row_len = (int)row_len;
    sl.nz0 = nz0;
    sl.nnz = nnz;
    sl.dia0 = dia0;
    sl.dia_len = st.kind == SK_DIA ? dia_len : 0;
    sl.t0 = t0i;
    sl.t1 = t1i;
    sl.need[0] = stage_doubles_of(st, 8);
    sl.need[1] = stage_doubles_of(st, 4);
    S.dirty = true;
    return slot;
}

// Stages the records of the live systems' tiles, marks each system's
// finisher, and copies them behind the work on the stream (pinned double
// buffer).
template <class T>
static void session_upload_active(Session& S) {
    Batch& B = S.B;
    const std::vector<TileVec<T>>& tv = host_records<T>(B);
    char* base = (char*)B.pin_act[B.pin_next];
    B.pin_next ^= 1;
    TileVec<T>* rec = (TileVec<T>*)base;
    Active A;
    int need = kStageMinDoubles;
    for (int k = 0; k < (int)S.slots.size(); ++k) {
        const SessSlot& sl = S.slots[k];
        if (!sl.used || !sl.live) continue;
        ++A.ns;
        need = std::max(need, sl.need[sizeof(T) == 4 ? 1 : 0]);
        for (int t = sl.t0; t < sl.t1; ++t) {
            rec[A.nt++] = tv[t];
            A.kinds |= kind_bit(tv[t].vec);
        }
    }
    mark_finishers(rec, A.nt, (int)S.slots.size());
    hipStream_t s = B.ctx->stream;
    if (A.nt > 0) {
        HSPS_HIP(hipMemcpyAsync((sizeof(T) == 4 ? B.tva32 : B.tva64).p, rec, sizeof(TileVec<T>) * (size_t)A.nt,
                                hipMemcpyHostToDevice, s));
    }
    S.act = A;
    S.stage = need;
}

// One chunk over the live systems; then the states of the chunk before are
// read and the slots whose solves ended since the last call are returned.
static std::vector<int> session_step(Session& S) {
    Batch& B = S.B;
    hipStream_t s = B.ctx->stream;
    const int chunk = std::max(1, std::min(S.p.check_every > 0 ? S.p.check_every : 32, 64));
    if (S.dirty) {
        if (S.prec == HSPS_MIXED) session_upload_active<float>(S);
        else session_upload_active<double>(S);
        S.dirty = false;
    }
    const int Sn = (int)S.slots.size();
    if (S.act.nt > 0) {
        BArgs a = args_of(B, S.p);
        a.stage_doubles = S.stage;
        const std::array<int, 4> key = {S.act.nt, S.act.ns, S.act.kinds, S.stage};
        int uses = 0;
        for (auto& kv : S.keys)
            if (kv.first == key) uses = ++kv.second;
        if (!uses) {
            if (S.keys.size() >= 64) S.keys.erase(S.keys.begin());
            S.keys.push_back({key, 1});
            uses = 1;
        }
        // capture a shape seen before; after repeated failed captures (other
        // threads of the process interfering) the session launches directly
        const bool graph = S.p.batch_graphs != 0 && uses >= 2 && B.capture_failures < 4;
        launch_chunk(B, a, 0, chunk, S.act, S.b.as<double>(), S.x.as<double>(), S.prec, s, graph);
    }
    BState* slotp = S.pin + (size_t)(S.seq & 1) * Sn;
    HSPS_HIP(hipMemcpyAsync(slotp, B.st.p, sizeof(BState) * (size_t)Sn, hipMemcpyDeviceToHost, s));
    HSPS_HIP(hipEventRecord(S.ev[S.seq & 1], s));
    const long long this_seq = S.seq++;
    std::vector<int> out;
    if (S.pending) {
        const long long prev_seq = this_seq - 1;
        HSPS_HIP(hipEventSynchronize(S.ev[prev_seq & 1]));
        const BState* prev = S.pin + (size_t)(prev_seq & 1) * Sn;
        for (int k = 0; k < Sn; ++k) {
            SessSlot& sl = S.slots[k];
            if (!sl.used || !sl.live || sl.join_seq > prev_seq) continue;
            if (prev[k].halt) {
                sl.live = 0;
                sl.done = 1;
                sl.conv = prev[k].conv;
                out.push_back(k);
                S.dirty = true;
            }
        }
    }
    S.pending = true;
    return out;
}

// The result of a slot whose solve ended: x += x_lo, the FP64 true residual,
// then the state and x (host and / or device copies); the slot is released.
static BState session_take(Session& S, int slot, double* h_x, double* d_x) {
    Batch& B = S.B;
    if (slot < 0 || slot >= (int)S.slots.size() || !S.slots[slot].used || !S.slots[slot].done)
        throw Error(HSPS_ERR_INVALID, "hsps_session_result: the slot holds no finished solve");
    SessSlot& sl = S.slots[slot];
    hipStream_t s = B.ctx->stream;
    const BArgs a = args_of(B, S.p);
    const int nt = sl.t1 - sl.t0;
    double* x = S.x.as<double>();
    if (S.prec == HSPS_MIXED)
        hipLaunchKernelGGL(k_b_flush_tiles, dim3(nt), dim3(kBlock), 0, s, a, sl.t0, x, B.xlo.as<float>());
    hipLaunchKernelGGL(k_b_final, dim3(nt), dim3(kBlock), 0, s, a, sl.t0, (const double*)S.b.as<double>(),
                       (const double*)x);
    hipLaunchKernelGGL((k_b_finish<1, OpFinal>), dim3(1), dim3(kFinishThreads), 0, s, a, slot, (const int*)nullptr,
                       OpFinal{});
    HSPS_CHECK_LAUNCH();
    BState st{};
    HSPS_HIP(hipMemcpyAsync(&st, B.st.as<BState>() + slot, sizeof(BState), hipMemcpyDeviceToHost, s));
    if (h_x) HSPS_HIP(hipMemcpyAsync(h_x, x + sl.row0, sizeof(double) * sl.n, hipMemcpyDeviceToHost, s));
    if (d_x) HSPS_HIP(hipMemcpyAsync(d_x, x + sl.row0, sizeof(double) * sl.n, hipMemcpyDeviceToDevice, s));
    HSPS_HIP(hipStreamSynchronize(s));
    S.rows.give(sl.row0, sl.row_len);
    S.nzs.give(sl.nz0, sl.nnz);
    S.tiles.give(sl.t0, sl.t1 - sl.t0);
    S.dias.give(sl.dia0, sl.dia_len);
    sl = SessSlot{};
    S.free_slots.push_back(slot);
    return st;
}

}  // namespace hsps_dev

using namespace hsps_dev;

struct hsps_batch_s : Batch {};

extern "C" {

int hsps_batch_create(hsps_ctx ctx, int S, const int* row_off, const int* rp, const int* ci, const double* v,
                      hsps_batch* out) {
    return guarded([&] {
        if (!out) throw Error(HSPS_ERR_INVALID, "hsps_batch_create: bad argument");
        auto* B = new hsps_batch_s;
        try {
            fill_batch(*B, ctx, S, row_off, rp, ci, v);
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
            const int tr = std::min(p.batch_tile_rows, 1 << 16);
            if (tr != B->tile_rows || tr != B->tile_rows_dia) build_tiles(*B, tr, tr);
        } else if (B->tile_rows == 0) {
            ensure_work(*B, p.precision == HSPS_MIXED ? HSPS_MIXED : HSPS_FP64);
            autotune(*B, b, x, p);
        }
        // HSPS_BATCH_NO_GRAPH=1 launches the iteration chunks directly (for
        // tracing tools that do not follow graph launches).
        static const bool no_graph = std::getenv("HSPS_BATCH_NO_GRAPH") != nullptr;
        std::vector<BState> st = run_batch(*B, b, x, p, !no_graph && p.batch_graphs != 0);
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

int hsps_batch_storage(hsps_batch B, int* n_csr, int* n_dia, int* n_stencil) {
    return guarded([&] {
        if (!B) throw Error(HSPS_ERR_INVALID, "hsps_batch_storage: null batch");
        if (n_csr) *n_csr = B->n_kind[SK_CSR];
        if (n_dia) *n_dia = B->n_kind[SK_DIA];
        if (n_stencil) *n_stencil = B->n_kind[SK_STENCIL];
        return HSPS_OK;
    });
}

}  // extern "C"

// ---- sessions: C ABI -------------------------------------------------------------------------
struct hsps_session_s : Session {
    std::vector<int> finished;  // slots whose end was observed, not yet handed out by step
};

extern "C" {

int hsps_session_create(hsps_ctx ctx, int max_systems, long long max_rows, long long max_nnz, const hsps_params* pp,
                        hsps_session* out) {
    return guarded([&] {
        if (!ctx || !out) throw Error(HSPS_ERR_INVALID, "hsps_session_create: null argument");
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        auto* S = new hsps_session_s;
        try {
            session_alloc(*S, ctx, max_systems, max_rows, max_nnz, p);
        } catch (...) {
            delete S;
            throw;
        }
        ctx_retain(ctx);
        *out = S;
        return HSPS_OK;
    });
}

int hsps_session_destroy(hsps_session S) {
    return guarded([&] {
        if (S) {
            Ctx* c = S->B.ctx;
            (void)hipSetDevice(c->device);
            (void)hipStreamSynchronize(c->stream);
            (void)hipStreamSynchronize(S->B.aux);
            delete S;
            ctx_release(c);
        }
        return HSPS_OK;
    });
}

int hsps_session_add(hsps_session S, int n, const int* rp, const int* ci, const double* v, const double* b,
                     const double* d_x0, int* slot) {
    return guarded([&] {
        if (!S || !slot) throw Error(HSPS_ERR_INVALID, "hsps_session_add: null argument");
        S->B.ctx->set_current();
        const int k = session_add(*S, n, rp, ci, v, b, d_x0);
        if (k < 0) throw Error(HSPS_ERR_NOMEM, "hsps_session_add: the session has no room for the system now");
        *slot = k;
        return HSPS_OK;
    });
}

int hsps_session_step(hsps_session S, int* done, int* converged, int cap, int* n_done) {
    return guarded([&] {
        if (!S || !n_done || (cap > 0 && !done)) throw Error(HSPS_ERR_INVALID, "hsps_session_step: null argument");
        S->B.ctx->set_current();
        for (int k : session_step(*S)) S->finished.push_back(k);
        const int m = std::min<int>(std::max(cap, 0), (int)S->finished.size());
        std::copy(S->finished.begin(), S->finished.begin() + m, done);
        if (converged)
            for (int i = 0; i < m; ++i) converged[i] = S->slots[done[i]].conv;
        S->finished.erase(S->finished.begin(), S->finished.begin() + m);
        *n_done = m;
        return HSPS_OK;
    });
}

int hsps_session_result(hsps_session S, int slot, double* h_x, double* d_x, int* iterations, int* converged,
                        double* true_res) {
    return guarded([&] {
        if (!S) throw Error(HSPS_ERR_INVALID, "hsps_session_result: null session");
        S->B.ctx->set_current();
        const BState st = session_take(*S, slot, h_x, d_x);
        if (iterations) *iterations = st.iter;
        if (converged) *converged = st.conv;
        if (true_res) *true_res = st.true_rel;
        return st.conv ? 1 : 0;
    });
}

int hsps_session_live(hsps_session S) {
    if (!S) return HSPS_ERR_INVALID;
    int live = 0;
    for (const auto& sl : S->slots) live += sl.used && sl.live;
    return live;
}

}  // extern "C"
