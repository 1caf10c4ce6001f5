// =============================================================================
// amg_setup.hip — the SA-AMG hierarchy built on the device (amg.cpp:183-237).
// Per level:
//   strength pattern |a_ij| >= theta * max_{l!=i}|a_il|     device (amg.cpp:39-46, 56)
//   greedy aggregation over that pattern                  host   (amg.cpp:48-60)
//   Jacobi-smoothed prolongation P, diagonal skipped      device (amg.cpp:101-178)
//   R = P^T                                               device (sparse_matrix.cpp:138-165)
//   Ac = R (A P), two sparse products                     device (sparse_matrix.cpp:171-237)
//   1 / diag(Ac)                                          device (amg.cpp:224-228)
// The aggregation stays on the host because its decisions depend on visiting
// rows in order; it reads only the strength pattern.  Every device stage
// yields the reference's values bit for bit: each entry is accumulated with
// the same operations in the same order (P: one rounded product and one
// subtraction per contributing entry, in row order; products: a fused
// multiply-add chain in (A-row, B-row) order), with the same drop rules and
// ascending columns.
//
// Sparse products run one 64-lane wave per output row.  The wave walks row i
// of A in order; for each entry a_ik its lanes take row k of B (distinct
// columns, so no two lanes touch one accumulator in a step) and add
// fma(a_ik, b_kl, acc_l) into an LDS accumulator keyed by column l:
//   * dense window  (B has <= kDenseCols columns, or a row too wide for the
//     hash): acc indexed by column; the row is emitted by scanning the window
//     in column order;
//   * hash          (otherwise): open-addressing table of 512 slots (rows of
//     <= 256 products) or 1024, a row that overflows it retried in 2048 slots;
//     kept entries are compacted and ranked.  The B-row loads of the next few
//     steps are in flight while one is accumulated.
// A count pass sizes each row (drop rule applied), a scan gives row pointers,
// and a fill pass recomputes the same accumulation and writes.
// =============================================================================
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <string>

#include "device_common.hpp"
#include "host_setup.hpp"
#include "precond.hpp"

namespace hsps_dev {
namespace {

constexpr double kEpsD = std::numeric_limits<double>::epsilon();
constexpr double kDropP = 1e-14;              // amg.cpp:160
constexpr double kDropGalerkin = kEpsD * kEpsD;  // sparse_matrix.cpp:209
constexpr int kDenseCols = 8192;              // dense LDS window: one FP64 accumulator per column (64 KB)
constexpr int kPThread = 12;                  // P rows from A rows up to this length: one thread each
constexpr int kPShort = 256;                  // longer ones up to this length: one wave each (k_p_wave)
constexpr int kStrongThread = 12;             // strength rows: one thread each up to this average length
constexpr int kPWindow = kDenseCols - 128;    // long P rows: window + 64-entry staging in 64 KB
constexpr int kHashSmall = 512;               // rows with <= 256 products (never overflows)
constexpr int kHashMid = 1024;                // rows with more (<= 512 distinct columns; 20 KB)
constexpr int kHashLarge = 2048;              // rows the mid table overflows (<= 1024 distinct; 40 KB)
constexpr int kSmallProducts = kHashSmall / 2;
constexpr int kGemmAhead = 8;                 // B-row loads in flight ahead of the accumulation
constexpr int kMaxWaves = 8192;               // persistent one-wave workgroups per launch

struct CsrRef {
    int rows = 0, cols = 0;
    long long nnz = 0;
    const int* rp = nullptr;
    const int* ci = nullptr;
    const double* v = nullptr;
};

__device__ __forceinline__ unsigned long long lanes_below(int lane) { return (1ull << lane) - 1ull; }

// Value of lane j (wave-uniform j) in every lane.
__device__ __forceinline__ int bcast(int x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ double bcast(double x, int j) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), j);
    return __hiloint2double(hi, lo);
}

// Appends i to list where pred holds: one atomic per wave.  Every lane of the
// wave calls it.
__device__ __forceinline__ void wave_append(bool pred, int i, int* list, int* n) {
    const unsigned long long m = __ballot(pred);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n, __popcll(m));
    base = __shfl(base, leader);
    if (pred) list[base + __popcll(m & lanes_below(lane))] = i;
}

// ---- strength pattern (amg.cpp:39-46, 56) --------------------------------------------
// row_max_i = max_{j != i} |a_ij| (std::max order), thr = theta * row_max_i;
// j != i is strong when |a_ij| >= thr.  srp == nullptr: count per row;
// otherwise write the strong columns of row i, in row order, at srp[i].
__global__ void k_strong(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                         const double* __restrict__ v, double theta, int* __restrict__ cnt,
                         const int* __restrict__ srp, int* __restrict__ sci) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k0 = rp[i], k1 = rp[i + 1];
        double m = 0.0;
        for (int k = k0; k < k1; ++k)
            if (ci[k] != i) {
                const double a = fabs(v[k]);
                m = m < a ? a : m;
            }
        const double thr = theta * m;
        if (!srp) {
            int c = 0;
            for (int k = k0; k < k1; ++k) c += (ci[k] != i && fabs(v[k]) >= thr);
            cnt[i] = c;
        } else {
            int o = srp[i];
            for (int k = k0; k < k1; ++k)
                if (ci[k] != i && fabs(v[k]) >= thr) sci[o++] = ci[k];
        }
    }
}

// The same, one wave per row (levels whose rows average more than
// kStrongThread entries): the row maximum is a wave reduction (max is exact
// in any order) and the strong columns are written in row order by ballot.
__global__ __launch_bounds__(64) void k_strong_wave(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const double* __restrict__ v, double theta,
                                                    int* __restrict__ cnt, const int* __restrict__ srp,
                                                    int* __restrict__ sci) {
    const int lane = threadIdx.x;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int k0 = rp[i], k1 = rp[i + 1];
        double m = 0.0;
        for (int k = k0 + lane; k < k1; k += 64)
            if (ci[k] != i) {
                const double a = fabs(v[k]);
                m = m < a ? a : m;
            }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(m, off);
            m = m < o ? o : m;
        }
        const double thr = theta * m;
        int o = srp ? srp[i] : 0;
        for (int kb = k0; kb < k1; kb += 64) {
            const int k = kb + lane;
            int j = -1;
            bool strong = false;
            if (k < k1) {
                j = ci[k];
                strong = j != i && fabs(v[k]) >= thr;
            }
            const unsigned long long b = __ballot(strong);
            if (srp && strong) sci[o + __popcll(b & lanes_below(lane))] = j;
            o += __popcll(b);
        }
        if (!srp && lane == 0) cnt[i] = o;
    }
}

// ---- smoothed prolongation (amg.cpp:101-178) -----------------------------------
// Row i of P: entry (agg_i, 1) minus, for each off-diagonal a_ij in row order,
// t = (omega / a_ii) * a_ij at column agg_j (nothing when |omega / a_ii| <
// eps); entries with |v| <= 1e-14 are dropped, columns ascending.
//
// Rows of A up to kPThread entries: one thread each; the row's list lives in
// scratch at rp[i] + i (at most row length + 1 entries) and is sorted by
// column at the end.  Rows up to kPShort entries: one wave each (k_p_wave),
// same scratch layout.  Longer rows, and wave rows with more than 64 distinct
// aggregate columns: k_p_long.
__device__ __forceinline__ double p_weight(int i, int k0, int k1, const int* ci, const double* v, double omega) {
    for (int k = k0; k < k1; ++k)
        if (ci[k] == i) return fabs(v[k]) > kEpsD ? omega / v[k] : 0.0;
    return 0.0;
}

__global__ void k_p_classify(int n, const int* __restrict__ rp, int* __restrict__ list, int* __restrict__ n_list,
                             int* __restrict__ wlist, int* __restrict__ n_wlist) {
    for (int b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        const int i = b + threadIdx.x;
        const int L = i < n ? rp[i + 1] - rp[i] : 0;
        wave_append(i < n && L > kPShort, i, list, n_list);
        wave_append(i < n && L > kPThread && L <= kPShort, i, wlist, n_wlist);
    }
}

__global__ void k_p_rows(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                         const int* __restrict__ agg, double omega, int* __restrict__ lc, double* __restrict__ lv,
                         int* __restrict__ len, int* __restrict__ cnt) {
#pragma clang fp contract(off)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k0 = rp[i], k1 = rp[i + 1];
        if (k1 - k0 > kPThread) continue;
        int* c_ = lc + k0 + i;
        double* v_ = lv + k0 + i;
        int L = 1;
        c_[0] = agg[i];
        v_[0] = 1.0;
        const double w = p_weight(i, k0, k1, ci, v, omega);
        if (!(fabs(w) < kEpsD)) {
            for (int k = k0; k < k1; ++k) {
                const int j = ci[k];
                if (j == i) continue;
                const int c = agg[j];
                const double t = w * v[k];
                int q = 0;
                while (q < L && c_[q] != c) ++q;
                if (q == L) {
                    c_[L] = c;
                    v_[L] = 0.0;
                    ++L;
                }
                v_[q] = v_[q] - t;
            }
        }
        for (int a = 1; a < L; ++a) {  // columns are distinct: insertion sort
            const int kc = c_[a];
            const double kv = v_[a];
            int b = a - 1;
            while (b >= 0 && c_[b] > kc) {
                c_[b + 1] = c_[b];
                v_[b + 1] = v_[b];
                --b;
            }
            c_[b + 1] = kc;
            v_[b + 1] = kv;
        }
        int kept = 0;
        for (int q = 0; q < L; ++q) kept += fabs(v_[q]) > kDropP;
        len[i] = L;
        cnt[i] = kept;
    }
}

// One wave per row.  Lane q owns the row's q-th distinct column (in order of
// first appearance, the reference's list order; lane 0: (agg_i, 1)) and its
// value.  For each 64-entry chunk the entries' (column, t) are broadcast in
// row order and each owner subtracts the t of its column: the same
// subtractions in the same order per column as the reference's list.  Owners
// are then ranked by column.  A row with more than 64 distinct columns is
// appended to `longs` (len 0) for k_p_long.
__global__ __launch_bounds__(64) void k_p_wave(const int* __restrict__ rows, int n_rows, const int* __restrict__ rp,
                                               const int* __restrict__ ci, const double* __restrict__ v,
                                               const int* __restrict__ agg, double omega, int* __restrict__ lc,
                                               double* __restrict__ lv, int* __restrict__ len, int* __restrict__ cnt,
                                               int* __restrict__ longs, int* __restrict__ n_long) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    for (int r = blockIdx.x; r < n_rows; r += gridDim.x) {
        const int i = rows[r];
        const int k0 = rp[i], k1 = rp[i + 1];
        // the first diagonal entry in row order
        double w = 0.0;
        for (int kb = k0; kb < k1; kb += 64) {
            const int k = kb + lane;
            const unsigned long long m = __ballot(k < k1 && ci[k] == i);
            if (m) {
                const int kd = kb + __ffsll((long long)m) - 1;
                w = fabs(v[kd]) > kEpsD ? omega / v[kd] : 0.0;
                break;
            }
        }
        int ucol = lane == 0 ? agg[i] : -1;
        double acc = lane == 0 ? 1.0 : 0.0;
        int nu = 1;
        bool over = false;
        if (!(fabs(w) < kEpsD)) {
            for (int kb = k0; kb < k1 && !over; kb += 64) {
                const int k = kb + lane;
                int c = -1;
                double t = 0.0;
                if (k < k1) {
                    const int j = ci[k];
                    if (j != i) {
                        c = agg[j];
                        t = w * v[k];
                    }
                }
                const int nk = min(64, k1 - kb);
                // new columns: not owned yet and first in this chunk
                bool fresh = c >= 0;
                for (int q = 0; q < 64; ++q) {
                    const int oc = __builtin_amdgcn_readlane(ucol, q);
                    const int cc = __builtin_amdgcn_readlane(c, q);
                    if (oc == c || (q < lane && cc == c)) fresh = false;
                }
                unsigned long long fm = __ballot(fresh);
                if (nu + __popcll(fm) > 64) {
                    over = true;
                    break;
                }
                while (fm) {
                    const int e = __ffsll((long long)fm) - 1;
                    fm &= fm - 1;
                    const int ce = __builtin_amdgcn_readlane(c, e);
                    if (lane == nu) ucol = ce;
                    ++nu;
                }
                for (int e = 0; e < nk; ++e) {
                    const int ce = __builtin_amdgcn_readlane(c, e);
                    const double te = bcast(t, e);
                    if (ce >= 0 && ucol == ce) acc = acc - te;
                }
            }
        }
        if (over) {
            if (lane == 0) {
                len[i] = 0;
                longs[atomicAdd(n_long, 1)] = i;
            }
            continue;
        }
        int rank = 0;
        for (int q = 0; q < nu; ++q) rank += __builtin_amdgcn_readlane(ucol, q) < ucol;
        const bool own = lane < nu;
        const unsigned long long km = __ballot(own && fabs(acc) > kDropP);
        if (own) {
            lc[k0 + i + rank] = ucol;
            lv[k0 + i + rank] = acc;
        }
        if (lane == 0) {
            len[i] = nu;
            cnt[i] = __popcll(km);
        }
    }
}

__global__ void k_p_fill(int n, const int* __restrict__ rp, const int* __restrict__ lc, const double* __restrict__ lv,
                         const int* __restrict__ len, const int* __restrict__ prp, int* __restrict__ pci,
                         double* __restrict__ pv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (rp[i + 1] - rp[i] > kPShort) continue;
        const int b = rp[i] + i;
        int o = prp[i];
        for (int q = 0; q < len[i]; ++q)
            if (fabs(lv[b + q]) > kDropP) {
                pci[o] = lc[b + q];
                pv[o] = lv[b + q];
                ++o;
            }
    }
}

// Long rows: one wave per row over windows of W aggregate columns (LDS
// accumulator, W doubles, then a 64-entry staging area).  Lanes stage 64
// entries at a time (column in the window, t); lane 0 applies them in row
// order; the window is then emitted in column order.  PASS 0 counts, PASS 1
// writes at prp[i].
template <int PASS>
__global__ __launch_bounds__(64) void k_p_long(const int* __restrict__ rows, int n_rows, const int* __restrict__ rp,
                                               const int* __restrict__ ci, const double* __restrict__ v,
                                               const int* __restrict__ agg, double omega, int nc, int W,
                                               int* __restrict__ cnt, const int* __restrict__ prp,
                                               int* __restrict__ pci, double* __restrict__ pv) {
#pragma clang fp contract(off)
    extern __shared__ double acc[];
    double* st_t = acc + W;
    int* st_c = reinterpret_cast<int*>(st_t + 64);
    const int lane = threadIdx.x;
    for (int c = lane; c < W; c += 64) acc[c] = 0.0;
    __syncthreads();
    for (int r = blockIdx.x; r < n_rows; r += gridDim.x) {
        const int i = rows[r];
        const int k0 = rp[i], k1 = rp[i + 1];
        const int ai = agg[i];
        // first diagonal entry, found a chunk at a time
        double w = 0.0;
        for (int kb = k0; kb < k1; kb += 64) {
            const int k = kb + lane;
            const unsigned long long m = __ballot(k < k1 && ci[k] == i);
            if (m) {
                const int kd = kb + __ffsll((long long)m) - 1;
                w = fabs(v[kd]) > kEpsD ? omega / v[kd] : 0.0;
                break;
            }
        }
        const bool off = !(fabs(w) < kEpsD);
        int u = 0;
        const int o = PASS == 1 ? prp[i] : 0;
        for (int c0 = 0; c0 < nc; c0 += W) {
            const int wn = min(W, nc - c0);
            if (lane == 0 && ai >= c0 && ai < c0 + wn) acc[ai - c0] = 1.0;
            __syncthreads();
            if (off) {
                for (int kb = k0; kb < k1; kb += 64) {
                    const int k = kb + lane;
                    int c = -1;
                    double t = 0.0;
                    if (k < k1) {
                        const int j = ci[k];
                        if (j != i) {
                            const int cc = agg[j] - c0;
                            if (cc >= 0 && cc < wn) {
                                c = cc;
                                t = w * v[k];
                            }
                        }
                    }
                    st_c[lane] = c;
                    st_t[lane] = t;
                    __syncthreads();
                    if (lane == 0) {
                        const int q1 = min(64, k1 - kb);
                        for (int q = 0; q < q1; ++q) {
                            const int cq = st_c[q];
                            if (cq >= 0) acc[cq] = acc[cq] - st_t[q];
                        }
                    }
                    __syncthreads();
                }
            }
            for (int cb = 0; cb < wn; cb += 64) {
                const int c = cb + lane;
                const bool in = c < wn;
                const bool keep = in && fabs(acc[c]) > kDropP;
                const unsigned long long m = __ballot(keep);
                if (PASS == 1 && keep) {
                    const int p = o + u + __popcll(m & lanes_below(lane));
                    pci[p] = c0 + c;
                    pv[p] = acc[c];
                }
                u += __popcll(m);
                if (in) acc[c] = 0.0;
            }
            __syncthreads();
        }
        if (PASS == 0 && lane == 0) cnt[i] = u;
    }
}

// ---- transpose (sparse_matrix.cpp:138-165) ---------------------------------------
// The reference scatters row by row, so each column's entries come out in row
// order: a stable radix sort of the entries by column (values = entry index)
// gives the same order, whatever the column populations.
__global__ void k_col_count(int nnz, const int* __restrict__ ci, int* __restrict__ cnt) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) atomicAdd(&cnt[ci[k]], 1);
}

__global__ void k_entry_rows(int rows, const int* __restrict__ rp, int* __restrict__ row_of, int* __restrict__ idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            row_of[k] = i;
            idx[k] = k;
        }
}

__global__ void k_t_gather(int nnz, const int* __restrict__ order, const int* __restrict__ row_of,
                           const double* __restrict__ v, int* __restrict__ tci, double* __restrict__ tv) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += gridDim.x * blockDim.x) {
        const int k = order[p];
        tci[p] = row_of[k];
        tv[p] = v[k];
    }
}

// ---- sparse product C = A B (sparse_matrix.cpp:171-237) --------------------------
// Rows travel as RowRef {i, A.rp[i], A.rp[i + 1], 0} (one 16-byte load), and
// each A entry p carries its B row's bounds bb[p] = {B.rp[k], B.rp[k + 1]}
// (k = A.ci[p]), so a row's first loads are two independent ones instead of a
// chain of four.
__global__ void k_gemm_bounds(long long nnz, const int* __restrict__ aci, const int* __restrict__ brp,
                              int2* __restrict__ bb) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nnz;
         p += (long long)gridDim.x * blockDim.x) {
        const int k = aci[p];
        bb[p] = make_int2(brp[k], brp[k + 1]);
    }
}

// Appends r to list where pred holds: one atomic per wave (wave_append for
// RowRefs).
__device__ __forceinline__ void wave_append_ref(bool pred, int4 r, int4* list, int* n) {
    const unsigned long long m = __ballot(pred);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n, __popcll(m));
    base = __shfl(base, leader);
    if (pred) list[base + __popcll(m & lanes_below(lane))] = r;
}

// Product count of each row, rows split into the small- and large-hash lists.
__global__ void k_gemm_classify(CsrRef A, const int2* __restrict__ bb, int4* __restrict__ list_small,
                                int4* __restrict__ list_large, int* __restrict__ n_lists) {
    for (int b = blockIdx.x * blockDim.x; b < A.rows; b += gridDim.x * blockDim.x) {
        const int i = b + threadIdx.x;
        const bool valid = i < A.rows;
        long long ub = 0;
        int a0 = 0, a1 = 0;
        if (valid) {
            a0 = A.rp[i];
            a1 = A.rp[i + 1];
            for (int pa = a0; pa < a1; ++pa) {
                const int2 e = bb[pa];
                ub += e.y - e.x;
            }
        }
        const int4 r = make_int4(i, a0, a1, 0);
        wave_append_ref(valid && ub <= kSmallProducts, r, list_small, &n_lists[0]);
        wave_append_ref(valid && ub > kSmallProducts, r, list_large, &n_lists[1]);
    }
}

// A 64-entry chunk of A's row starting at p0 (row end a1): lane e holds entry
// p0 + e's value and B-row bounds.  Every lane loads (a lane past the row
// loads entry 0, A having entries, and is never read), so no load sits under
// a branch.
__device__ __forceinline__ void load_chunk(const CsrRef& A, const int2* bb, int p0, int a1, double& av, int2& b) {
    const int p = p0 + (int)(threadIdx.x & 63);
    const int pc = p < a1 ? p : 0;
    av = A.v[pc];
    b = bb[pc];
}

// The rows of the list, one wave each (r = blockIdx.x, += gridDim.x).  While a
// row runs, the next row's RowRef and first A chunk and the RowRef after it
// are in flight.  body(row, av, b) gets the row's first chunk.
template <class Body>
__device__ __forceinline__ void for_each_row(const int4* rows, int n_rows, const CsrRef& A, const int2* bb,
                                             Body body) {
    int r = blockIdx.x;
    if (r >= n_rows) return;
    const int g = gridDim.x;
    int4 cur = rows[r];
    double av;
    int2 b;
    load_chunk(A, bb, cur.y, cur.z, av, b);
    int4 nxt = rows[min(r + g, n_rows - 1)];
    for (; r < n_rows; r += g) {
        double nav;
        int2 nb;
        load_chunk(A, bb, nxt.y, nxt.z, nav, nb);
        const int4 nxt2 = rows[min(r + 2 * g, n_rows - 1)];
        body(cur, av, b);
        cur = nxt;
        av = nav;
        b = nb;
        nxt = nxt2;
    }
}

// The steps of one 64-entry chunk of A's row i (lane e holds entry e: its
// B-row bounds [b0, b1)): a step is one 64-entry run of one B row, in A's
// order, runs in B's order; an empty B row is one empty step.  f(j, l, v)
// consumes a step in every lane: j = the chunk entry, (l, v) this lane's B
// entry (l = -1 past the run).  The B entries of the next D steps are loaded
// while one is consumed, so the consumption order is the product order and
// global-load latency overlaps the accumulation.  Every lane loads in every
// step (a lane past its run, or a step past the chunk, loads B entry 0 and
// ignores it), so no load sits under a branch and the compiler's wait counts
// stay exact.  Stops early once stop is set (by f).
template <int D, class F>
__device__ __forceinline__ void for_each_step(const CsrRef& B, int b0, int b1, int na, const bool& stop, F f) {
    constexpr int safe = 0;  // callers launch only when B has entries
    const int lane = threadIdx.x & 63;
    int pj = 0, pp = bcast(b0, 0), pe = bcast(b1, 0);  // producer cursor (wave-uniform)
    int rl[D], rj[D];
    double rv[D];
    bool rin[D];
    auto issue = [&](int d) {
        const bool live = pj < na;
        const int p = pp + lane;
        rin[d] = live && p < pe;
        const int pc = rin[d] ? p : safe;
        rl[d] = B.ci[pc];
        rv[d] = B.v[pc];
        rj[d] = live ? pj : -1;
        if (live) {
            pp += 64;
            if (pp >= pe && ++pj < na) {
                pp = bcast(b0, pj);
                pe = bcast(b1, pj);
            }
        }
    };
#pragma unroll
    for (int d = 0; d < D; ++d) issue(d);
    while (rj[0] >= 0 && !stop) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (rj[d] >= 0 && !stop) f(rj[d], rin[d] ? rl[d] : -1, rv[d]);
            issue(d);
        }
    }
}

// The chunks of a row (the first one preloaded), each through for_each_step.
template <class F>
__device__ __forceinline__ void for_each_product(const CsrRef& A, const int2* bb, const CsrRef& B, int4 row,
                                                 double av, int2 b, const bool& stop, F f) {
    const int a1 = row.z;
    for (int base = row.y; base < a1 && !stop; base += 64) {
        if (base != row.y) load_chunk(A, bb, base, a1, av, b);
        const int na = min(64, a1 - base);
        for_each_step<kGemmAhead>(B, b.x, b.y, na, stop, [&](int j, int l, double v) { f(bcast(av, j), l, v); });
    }
}

// One wave per row, open-addressing hash of CAP slots in LDS.  PASS 0 writes
// the kept count (rows whose distinct columns exceed CAP / 2 are appended to
// `spill` and left to the next larger table or the dense-window kernel); PASS 1
// writes the row at crp[i] in ascending columns (and skips rows that spilled).
// Occupancy is checked after every step of <= 64 insertions, so the table
// never exceeds CAP / 2 + 64 slots and probing ends.  Occupied slots are
// listed as they fill, so emitting and clearing a row visits only those.
// Products are accumulated in the row's order (for_each_product).
template <int CAP, int PASS>
__global__ __launch_bounds__(64) void k_gemm_hash(const int4* __restrict__ rows, int n_rows, CsrRef A,
                                                  const int2* __restrict__ bb, CsrRef B, int* __restrict__ cnt,
                                                  const int* __restrict__ crp, int* __restrict__ cci,
                                                  double* __restrict__ cv, int4* __restrict__ spill,
                                                  int* __restrict__ n_spill) {
    constexpr int HALF = CAP / 2;
    constexpr int BITS = __builtin_ctz(CAP);
    __shared__ int key[CAP];
    __shared__ double val[CAP];
    __shared__ int occl[HALF + 64];
    __shared__ int lk[PASS == 1 ? HALF : 1];
    __shared__ double lvv[PASS == 1 ? HALF : 1];
    const int lane = threadIdx.x;
    for (int s = lane; s < CAP; s += 64) {
        key[s] = -1;
        val[s] = 0.0;
    }
    __syncthreads();
    for_each_row(rows, n_rows, A, bb, [&](int4 row, double av, int2 b) {
        const int i = row.x;
        int occ = 0;
        bool over = false;
        // insert (l, av * bv) from this lane when l >= 0; over once the table is half full
        for_each_product(A, bb, B, row, av, b, over, [&](double avj, int l, double bv) {
            bool fresh = false;
            int s = 0;
            if (l >= 0) {
                s = (int)(((unsigned)l * 2654435761u) >> (32 - BITS));
                while (true) {
                    const int old = atomicCAS(&key[s], -1, l);
                    if (old == -1) { fresh = true; break; }
                    if (old == l) break;
                    s = (s + 1) & (CAP - 1);
                }
                val[s] = fma(avj, bv, val[s]);
            }
            const unsigned long long m = __ballot(fresh);
            if (fresh) occl[occ + __popcll(m & lanes_below(lane))] = s;
            occ += __popcll(m);
            if (occ > HALF) over = true;
        });
        __syncthreads();
        int u = 0;
        for (int e0 = 0; e0 < occ; e0 += 64) {
            const int e = e0 + lane;
            bool keep = false;
            int kk = 0;
            double vv = 0.0;
            if (e < occ) {
                const int s = occl[e];
                kk = key[s];
                vv = val[s];
                keep = !over && fabs(vv) > kDropGalerkin;
                key[s] = -1;
                val[s] = 0.0;
            }
            const unsigned long long m = __ballot(keep);
            if (PASS == 1 && keep) {
                const int pos = u + __popcll(m & lanes_below(lane));
                lk[pos] = kk;
                lvv[pos] = vv;
            }
            u += __popcll(m);
        }
        __syncthreads();
        if (PASS == 0) {
            if (lane == 0) {
                cnt[i] = over ? 0 : u;
                if (over) spill[atomicAdd(n_spill, 1)] = row;
            }
        } else if (!over) {
            const int o = crp[i];
            for (int e = lane; e < u; e += 64) {
                const int kk = lk[e];
                int rank = 0;
                for (int f = 0; f < u; ++f) rank += lk[f] < kk;
                cci[o + rank] = kk;
                cv[o + rank] = lvv[e];
            }
        }
        __syncthreads();
    });
}

// One wave per row, dense accumulator over a window of W columns in LDS
// (W doubles, dynamic), products in the row's order (for_each_product).
// Rows wider than the window are processed window by window, each a full
// pass over the row's products.  An untouched column holds 0 and falls under
// the drop rule, so no touched flags are kept.
template <int PASS>
__global__ __launch_bounds__(64) void k_gemm_dense(const int4* __restrict__ rows, int n_rows, CsrRef A,
                                                   const int2* __restrict__ bb, CsrRef B, int W,
                                                   int* __restrict__ cnt, const int* __restrict__ crp,
                                                   int* __restrict__ cci, double* __restrict__ cv) {
    extern __shared__ double acc[];
    const int lane = threadIdx.x;
    for (int c = lane; c < W; c += 64) acc[c] = 0.0;
    __syncthreads();
    const bool never = false;
    for_each_row(rows, n_rows, A, bb, [&](int4 row, double av, int2 b) {
        const int i = row.x;
        int u = 0;
        const int o = PASS == 1 ? crp[i] : 0;
        for (int c0 = 0; c0 < B.cols; c0 += W) {
            const int wn = min(W, B.cols - c0);
            for_each_product(A, bb, B, row, av, b, never, [&](double avj, int l, double v) {
                l -= c0;
                if (l >= 0 && l < wn) acc[l] = fma(avj, v, acc[l]);
            });
            __syncthreads();
            for (int cb = 0; cb < wn; cb += 64) {
                const int c = cb + lane;
                const bool in = c < wn;
                const bool keep = in && fabs(acc[c]) > kDropGalerkin;
                const unsigned long long m = __ballot(keep);
                if (PASS == 1 && keep) {
                    const int p = o + u + __popcll(m & lanes_below(lane));
                    cci[p] = c0 + c;
                    cv[p] = acc[c];
                }
                u += __popcll(m);
                if (in) acc[c] = 0.0;
            }
            __syncthreads();
        }
        if (PASS == 0 && lane == 0) cnt[i] = u;
    });
}

// ---- 1 / diagonal (amg.cpp:189-196, 224-228) and the raw diagonal --------------
__global__ void k_inv_diag(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                           double* __restrict__ d64, float* __restrict__ d32) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                a = v[k];
                break;
            }
        const double d = fabs(a) > kEpsD ? 1.0 / a : 0.0;
        d64[i] = d;
        d32[i] = (float)d;
    }
}

// ---- host-side helpers ------------------------------------------------------------
inline int blocks_for(long n, int threads = kBlock) {
    long g = (n + threads - 1) / threads;
    return (int)std::max(1L, std::min(g, 65536L));
}

// rp[0] = 0, rp[1..n] = inclusive scan of cnt; rp also copied to the host.
// Throws when the total leaves the 32-bit index range.
void scan_rows(Ctx* c, const int* cnt, int n, DevBuf& rp, std::vector<int>& h_rp, hipStream_t s, const char* what) {
    rp = DevBuf(c, (size_t)(n + 1) * sizeof(int));
    HSPS_HIP(hipMemsetAsync(rp.p, 0, sizeof(int), s));
    if (n > 0) {
        size_t tb = 0;
        HSPS_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, cnt, rp.as<int>() + 1, n, s));
        DevBuf tmp(c, std::max<size_t>(tb, 16));
        HSPS_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, cnt, rp.as<int>() + 1, n, s));
        h_rp.resize(n + 1);
        HSPS_HIP(hipMemcpyAsync(h_rp.data(), rp.p, (size_t)(n + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
    } else {
        h_rp.assign(1, 0);
    }
    for (int i = 0; i < n; ++i)
        if (h_rp[i + 1] < h_rp[i])
            throw Error(HSPS_ERR_NOMEM, std::string("AMG setup: ") + what + " exceeds 32-bit indexing");
}

CsrRef ref_of(const Csr& A) {
    CsrRef r;
    r.rows = A.rows;
    r.cols = A.cols;
    r.nnz = A.nnz;
    r.rp = A.rp.as<int>();
    r.ci = A.ci.as<int>();
    r.v = A.v64.as<double>();
    return r;
}

// C = A B with the reference's accumulation order, drop rule and column order.
std::unique_ptr<Csr> spgemm_dev(Ctx* c, const CsrRef& A, const CsrRef& B, bool final_level, hipStream_t s) {
    const int m = A.rows;
    DevBuf cnt(c, (size_t)std::max(m, 1) * sizeof(int));
    const bool dense = B.cols <= kDenseCols;
    const int W = std::min(std::max(B.cols, 1), kDenseCols);
    const size_t dense_lds = (size_t)W * sizeof(double);
    const bool any = m > 0 && A.nnz > 0 && B.nnz > 0;  // else every row is empty
    DevBuf lists, counters, spill, bbuf;
    int n_small = 0, n_large = 0, n_spill = 0, n_spill2 = 0;
    if (any) {
        bbuf = DevBuf(c, (size_t)A.nnz * sizeof(int2));
        hipLaunchKernelGGL(k_gemm_bounds, dim3(blocks_for((int)std::min<long long>(A.nnz, 1 << 30))), dim3(kBlock), 0, s,
                           A.nnz, A.ci, B.rp, bbuf.as<int2>());
        HSPS_CHECK_LAUNCH();
        lists = DevBuf(c, (size_t)m * 2 * sizeof(int4));
        spill = DevBuf(c, (size_t)m * 2 * sizeof(int4));
        counters = DevBuf(c, 4 * sizeof(int));
        HSPS_HIP(hipMemsetAsync(counters.p, 0, 4 * sizeof(int), s));
        hipLaunchKernelGGL(k_gemm_classify, dim3(blocks_for(m)), dim3(kBlock), 0, s, A, bbuf.as<int2>(),
                           lists.as<int4>(), lists.as<int4>() + m, counters.as<int>());
        HSPS_CHECK_LAUNCH();
        int h[2];
        HSPS_HIP(hipMemcpyAsync(h, counters.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
        n_small = h[0];
        n_large = h[1];
    }
    const int2* bb = bbuf.as<int2>();
    const int4* small = lists.as<int4>();
    const int4* large = lists.as<int4>() + m;
    int4* spill1 = spill.as<int4>();
    int4* spill2 = spill.as<int4>() + m;
    // dense: both lists through the window kernel.  Otherwise the small
    // table for small rows, the mid table for large ones; rows the mid table
    // overflows (spill1, counters[2]) go to the large table, rows that one
    // overflows (spill2, counters[3]) to the window kernel.
    auto run = [&](auto pass_tag, const DevBuf* crp, int* cci, double* cv) {
        constexpr int PASS = decltype(pass_tag)::value;
        if (!any) {
            if (PASS == 0 && m > 0) HSPS_HIP(hipMemsetAsync(cnt.p, 0, (size_t)m * sizeof(int), s));
            return;
        }
        int* out_rp = crp ? crp->as<int>() : nullptr;
        auto window = [&](const int4* list, int nl) {
            if (nl)
                hipLaunchKernelGGL(k_gemm_dense<PASS>, dim3(std::min(nl, kMaxWaves)), dim3(64), dense_lds, s, list, nl,
                                   A, bb, B, W, cnt.as<int>(), out_rp, cci, cv);
            HSPS_CHECK_LAUNCH();
        };
        if (dense) {
            window(small, n_small);
            window(large, n_large);
            return;
        }
        int* ns1 = counters.as<int>() + 2;
        int* ns2 = counters.as<int>() + 3;
        if (n_small)
            hipLaunchKernelGGL((k_gemm_hash<kHashSmall, PASS>), dim3(std::min(n_small, kMaxWaves)), dim3(64), 0, s,
                               small, n_small, A, bb, B, cnt.as<int>(), out_rp, cci, cv, spill1, ns1);
        HSPS_CHECK_LAUNCH();
        if (n_large)
            hipLaunchKernelGGL((k_gemm_hash<kHashMid, PASS>), dim3(std::min(n_large, kMaxWaves)), dim3(64), 0, s,
                               large, n_large, A, bb, B, cnt.as<int>(), out_rp, cci, cv, spill1, ns1);
        HSPS_CHECK_LAUNCH();
        if (PASS == 0) {
            HSPS_HIP(hipMemcpyAsync(&n_spill, ns1, sizeof(int), hipMemcpyDeviceToHost, s));
            HSPS_HIP(hipStreamSynchronize(s));
        }
        if (n_spill)
            hipLaunchKernelGGL((k_gemm_hash<kHashLarge, PASS>), dim3(std::min(n_spill, kMaxWaves)), dim3(64), 0, s,
                               spill1, n_spill, A, bb, B, cnt.as<int>(), out_rp, cci, cv, spill2, ns2);
        HSPS_CHECK_LAUNCH();
        if (PASS == 0 && n_spill) {
            HSPS_HIP(hipMemcpyAsync(&n_spill2, ns2, sizeof(int), hipMemcpyDeviceToHost, s));
            HSPS_HIP(hipStreamSynchronize(s));
        }
        window(spill2, n_spill2);
    };
    run(std::integral_constant<int, 0>{}, nullptr, nullptr, nullptr);
    DevBuf crp;
    std::vector<int> h_rp;
    scan_rows(c, cnt.as<int>(), m, crp, h_rp, s, "sparse product");
    const int nnz = h_rp[m];
    DevBuf cci(c, (size_t)nnz * sizeof(int)), cv(c, (size_t)nnz * sizeof(double));
    run(std::integral_constant<int, 1>{}, &crp, cci.as<int>(), cv.as<double>());
    return csr_adopt(c, m, B.cols, std::move(crp), std::move(cci), std::move(cv), std::move(h_rp), final_level);
}

}  // namespace

// The hierarchy for M (kind AMG) over A; see the file comment.
void amg_setup_device(Precond& M, const Csr& A, const hsps_params& p, AmgSetupTimes* times) {
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    const int coarse_limit = 50;  // amg.hpp:99
    Ctx* c = M.ctx;
    hipStream_t s = c->stream;
    c->set_current();
    M.amg.clear();
    auto inv_diag = [&](AmgLevelDev& L, const Csr& F) {
        L.dinv64 = DevBuf(c, (size_t)std::max(F.rows, 1) * sizeof(double));
        L.dinv32 = DevBuf(c, (size_t)std::max(F.rows, 1) * sizeof(float));
        if (F.rows)
            hipLaunchKernelGGL(k_inv_diag, dim3(blocks_for(F.rows)), dim3(kBlock), 0, s, F.rows, F.rp.as<int>(),
                               F.ci.as<int>(), F.v64.as<double>(), L.dinv64.as<double>(), L.dinv32.as<float>());
        HSPS_CHECK_LAUNCH();
    };
    M.amg.emplace_back();
    inv_diag(M.amg[0], A);
    const Csr* F = &A;
    for (int lv = 0; lv < p.amg_levels - 1; ++lv) {  // amg.cpp:198-234
        const int n = F->rows;
        if (n <= coarse_limit) break;
        // strength pattern -> host greedy aggregation
        auto t0 = clk::now();
        DevBuf cnt(c, (size_t)n * sizeof(int)), srp;
        std::vector<int> h_srp;
        const bool strong_wave = F->nnz > (long long)kStrongThread * n;
        auto strong = [&](int* cnt_out, const int* srp_in, int* sci_out) {
            if (strong_wave)
                hipLaunchKernelGGL(k_strong_wave, dim3(std::min(n, kMaxWaves)), dim3(64), 0, s, n, F->rp.as<int>(),
                                   F->ci.as<int>(), F->v64.as<double>(), p.amg_strong, cnt_out, srp_in, sci_out);
            else
                hipLaunchKernelGGL(k_strong, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, F->rp.as<int>(),
                                   F->ci.as<int>(), F->v64.as<double>(), p.amg_strong, cnt_out, srp_in, sci_out);
            HSPS_CHECK_LAUNCH();
        };
        strong(cnt.as<int>(), nullptr, nullptr);
        scan_rows(c, cnt.as<int>(), n, srp, h_srp, s, "strength pattern");
        std::vector<int> h_sci(h_srp[n]);
        {
            DevBuf sci(c, (size_t)std::max(h_srp[n], 1) * sizeof(int));
            strong(nullptr, srp.as<int>(), sci.as<int>());
            if (h_srp[n])
                HSPS_HIP(hipMemcpyAsync(h_sci.data(), sci.p, (size_t)h_srp[n] * sizeof(int), hipMemcpyDeviceToHost, s));
            HSPS_HIP(hipStreamSynchronize(s));
        }
        if (times) times->strength += secs(t0);
        t0 = clk::now();
        int nc = 0;
        std::vector<int> agg = hsps_host::aggregate_strong(n, h_srp, h_sci, &nc);
        if (times) times->aggregate += secs(t0);
        if (nc >= n) break;
        t0 = clk::now();
        DevBuf dagg(c, (size_t)n * sizeof(int));
        HSPS_HIP(hipMemcpyAsync(dagg.p, agg.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
        // P: short rows one thread each, medium rows one wave each (owner
        // lanes), long rows one wave each over LDS windows
        std::unique_ptr<Csr> P;
        {
            const size_t slots = (size_t)F->nnz + n;
            DevBuf lc(c, slots * sizeof(int)), lvals(c, slots * sizeof(double)), len(c, (size_t)n * sizeof(int));
            DevBuf longs(c, (size_t)n * sizeof(int)), waves(c, (size_t)n * sizeof(int)),
                n_lists(c, 2 * sizeof(int));
            HSPS_HIP(hipMemsetAsync(n_lists.p, 0, 2 * sizeof(int), s));
            hipLaunchKernelGGL(k_p_classify, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, F->rp.as<int>(),
                               longs.as<int>(), n_lists.as<int>(), waves.as<int>(), n_lists.as<int>() + 1);
            HSPS_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_p_rows, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, F->rp.as<int>(), F->ci.as<int>(),
                               F->v64.as<double>(), dagg.as<int>(), M.amg_omega, lc.as<int>(), lvals.as<double>(),
                               len.as<int>(), cnt.as<int>());
            HSPS_CHECK_LAUNCH();
            int h_n[2] = {0, 0};
            HSPS_HIP(hipMemcpyAsync(h_n, n_lists.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            HSPS_HIP(hipStreamSynchronize(s));
            if (h_n[1]) {
                hipLaunchKernelGGL(k_p_wave, dim3(std::min(h_n[1], kMaxWaves)), dim3(64), 0, s, waves.as<int>(), h_n[1],
                                   F->rp.as<int>(), F->ci.as<int>(), F->v64.as<double>(), dagg.as<int>(), M.amg_omega,
                                   lc.as<int>(), lvals.as<double>(), len.as<int>(), cnt.as<int>(), longs.as<int>(),
                                   n_lists.as<int>());
                HSPS_CHECK_LAUNCH();
                HSPS_HIP(hipMemcpyAsync(h_n, n_lists.p, sizeof(int), hipMemcpyDeviceToHost, s));
                HSPS_HIP(hipStreamSynchronize(s));
            }
            const int n_long = h_n[0];
            const int Wp = std::min(std::max(nc, 1), kPWindow);
            const size_t long_lds = (size_t)Wp * sizeof(double) + 64 * (sizeof(double) + sizeof(int));
            if (n_long)
                hipLaunchKernelGGL(k_p_long<0>, dim3(std::min(n_long, kMaxWaves)), dim3(64), long_lds, s,
                                   longs.as<int>(), n_long, F->rp.as<int>(), F->ci.as<int>(), F->v64.as<double>(),
                                   dagg.as<int>(), M.amg_omega, nc, Wp, cnt.as<int>(), nullptr, nullptr, nullptr);
            HSPS_CHECK_LAUNCH();
            DevBuf prp;
            std::vector<int> h_prp;
            scan_rows(c, cnt.as<int>(), n, prp, h_prp, s, "prolongation");
            const int pnnz = h_prp[n];
            DevBuf pci(c, (size_t)pnnz * sizeof(int)), pv(c, (size_t)pnnz * sizeof(double));
            hipLaunchKernelGGL(k_p_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, F->rp.as<int>(), lc.as<int>(),
                               lvals.as<double>(), len.as<int>(), prp.as<int>(), pci.as<int>(), pv.as<double>());
            HSPS_CHECK_LAUNCH();
            if (n_long)
                hipLaunchKernelGGL(k_p_long<1>, dim3(std::min(n_long, kMaxWaves)), dim3(64), long_lds, s,
                                   longs.as<int>(), n_long, F->rp.as<int>(), F->ci.as<int>(), F->v64.as<double>(),
                                   dagg.as<int>(), M.amg_omega, nc, Wp, cnt.as<int>(), prp.as<int>(), pci.as<int>(),
                                   pv.as<double>());
            HSPS_CHECK_LAUNCH();
            P = csr_adopt(c, n, nc, std::move(prp), std::move(pci), std::move(pv), std::move(h_prp), true);
        }
        if (times) times->prolongation += secs(t0);
        // R = P^T
        t0 = clk::now();
        std::unique_ptr<Csr> R;
        {
            const int pnnz = P->nnz;
            DevBuf ccnt(c, (size_t)nc * sizeof(int));
            HSPS_HIP(hipMemsetAsync(ccnt.p, 0, (size_t)nc * sizeof(int), s));
            if (pnnz)
                hipLaunchKernelGGL(k_col_count, dim3(blocks_for(pnnz)), dim3(kBlock), 0, s, pnnz, P->ci.as<int>(),
                                   ccnt.as<int>());
            HSPS_CHECK_LAUNCH();
            DevBuf rrp;
            std::vector<int> h_rrp;
            scan_rows(c, ccnt.as<int>(), nc, rrp, h_rrp, s, "restriction");
            DevBuf rci(c, (size_t)pnnz * sizeof(int)), rv(c, (size_t)pnnz * sizeof(double));
            if (pnnz) {
                DevBuf row_of(c, (size_t)pnnz * sizeof(int)), idx(c, (size_t)pnnz * sizeof(int));
                DevBuf keys_out(c, (size_t)pnnz * sizeof(int)), order(c, (size_t)pnnz * sizeof(int));
                hipLaunchKernelGGL(k_entry_rows, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, P->rp.as<int>(),
                                   row_of.as<int>(), idx.as<int>());
                HSPS_CHECK_LAUNCH();
                int bits = 1;
                while (bits < 31 && (1 << bits) < nc) ++bits;
                size_t tb = 0;
                HSPS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, P->ci.as<int>(), keys_out.as<int>(),
                                                            idx.as<int>(), order.as<int>(), pnnz, 0, bits, s));
                DevBuf tmp(c, std::max<size_t>(tb, 16));
                HSPS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, P->ci.as<int>(), keys_out.as<int>(),
                                                            idx.as<int>(), order.as<int>(), pnnz, 0, bits, s));
                hipLaunchKernelGGL(k_t_gather, dim3(blocks_for(pnnz)), dim3(kBlock), 0, s, pnnz, order.as<int>(),
                                   row_of.as<int>(), P->v64.as<double>(), rci.as<int>(), rv.as<double>());
                HSPS_CHECK_LAUNCH();
                HSPS_HIP(hipStreamSynchronize(s));
            }
            R = csr_adopt(c, nc, n, std::move(rrp), std::move(rci), std::move(rv), std::move(h_rrp), true);
        }
        if (times) times->transpose += secs(t0);
        // Ac = R (A P)
        t0 = clk::now();
        std::unique_ptr<Csr> AP = spgemm_dev(c, ref_of(*F), ref_of(*P), false, s);
        std::unique_ptr<Csr> Ac = spgemm_dev(c, ref_of(*R), ref_of(*AP), true, s);
        AP.reset();
        if (times) times->galerkin += secs(t0);
        AmgLevelDev& cur = M.amg.back();
        cur.agg = std::move(agg);
        cur.P = std::move(P);
        cur.R = std::move(R);
        M.amg.emplace_back();
        inv_diag(M.amg.back(), *Ac);
        M.amg.back().A = std::move(Ac);
        F = M.amg.back().A.get();
    }
    HSPS_HIP(hipStreamSynchronize(s));
}

}  // namespace hsps_dev
