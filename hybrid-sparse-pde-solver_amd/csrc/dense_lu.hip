// =============================================================================
// dense_lu.hip — the coarsest AMG level's dense solve operator, built on the
// device: G = A^-1 in the sense of the reference's direct_solve
// (amg.cpp:260-291: Gaussian elimination with partial pivoting, the first
// maximum |a_rc| as pivot, elimination of a column skipped when its pivot is
// below 1e-14, and x_i = 0 in back substitution when |u_ii| <= 1e-14).
//
// The matrix is scattered into a dense row-major N x N array (N = n rounded
// up to 64, identity on the padding), factored as P A = L U by a blocked
// right-looking LU (16-column panels), and G = U^-1 L^-1 P is formed by
// blocked triangular solves (64-row blocks) on the permuted identity.  The
// GEMM-shaped updates (the trailing matrix, and the off-diagonal blocks of
// both triangular solves) run on FP64 MFMA (v_mfma_f64_16x16x4_f64); the
// panel (pivot search, swaps, multipliers: one workgroup, the panel held in
// registers) and the diagonal blocks (one thread per column) are
// latency-bound.
// =============================================================================
#include <cmath>
#include <utility>

#include "dense_lu.hpp"
#include "device_common.hpp"

namespace hsps_dev {

constexpr int kLuNB = 64;           // padding multiple and block size of the inverse's triangular solves
constexpr int kPW = 16;             // LU panel width (the trailing GEMM's depth)
constexpr int kPanelThreads = 1024;
constexpr int kPanelRegRows = 2;    // panel rows per thread: the register panel takes R <= 2048
constexpr double kPivTiny = 1e-14;  // amg.cpp:275, 288

typedef double v4d __attribute__((ext_vector_type(4)));

// D = identity on the padding, zero elsewhere; then the CSR entries scattered
// in (amg.cpp:262-267 builds the same dense copy).
__global__ void k_dense_init(int N, int n, double* __restrict__ D) {
    const size_t total = (size_t)N * N;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / N), j = (int)(e % N);
        D[e] = (i == j && i >= n) ? 1.0 : 0.0;
    }
}
// One wave per row (the coarsest operators are nearly dense: ~860 entries a row at 2048^2).
__global__ void k_dense_scatter(int n, int N, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ v, double* __restrict__ D) {
    const int i = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (i >= n) return;
    for (int k = rp[i] + (threadIdx.x & 63); k < rp[i + 1]; k += 64) D[(size_t)i * N + ci[k]] = v[k];
}
__global__ void k_iota(int N, int* __restrict__ sigma) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) sigma[i] = i;
}

template <int... I, class F>
__device__ __forceinline__ void static_for(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Pivot of column cc among rows [cc, R) of a panel: the first row of maximal
// |a| (amg.cpp:271-273).  Each thread passes its best candidate (|a| = -1 for
// none; it scans its rows in increasing order, so a strict > keeps the first
// max); the winner is valid in every thread after the call (two barriers).
__device__ __forceinline__ int panel_pivot(double best, int bi, double* s_val, int* s_idx, int* s_piv) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    auto reduce = [&](double& b, int& i) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(b, o, 64);
            const int oi = __shfl_xor(i, o, 64);
            if (ob > b || (ob == b && oi < i)) { b = ob; i = oi; }
        }
    };
    reduce(best, bi);
    if (lane == 0) { s_val[wid] = best; s_idx[wid] = bi; }
    __syncthreads();
    if (wid == 0) {
        double b = lane < kPanelThreads / 64 ? s_val[lane] : -2.0;
        int i = lane < kPanelThreads / 64 ? s_idx[lane] : 0x7fffffff;
        reduce(b, i);
        if (lane == 0) *s_piv = i;
    }
    __syncthreads();
    return *s_piv;
}

// Panel factorisation of columns [k0, k0 + 16), rows [k0, N), one workgroup
// with the panel in registers (thread t holds local rows t and t + 1024, so
// R = N - k0 <= 2048).  Per column c: the pivot row and row c trade places
// through LDS; multipliers l_r = a_rc * (1 / a_cc), or 0 when |a_cc| < 1e-14
// (the column is not eliminated, amg.cpp:275); the rank-1 update covers the
// panel's remaining columns (the trailing GEMM updates the rest).  The row
// swaps outside the panel's columns are left to k_lu_laswp: the panel's net
// permutation goes to `swaps` as (count, then destination / source row pairs).
__global__ void __launch_bounds__(kPanelThreads)
k_lu_panel_reg(double* __restrict__ A, int N, int k0, int* __restrict__ piv, int* __restrict__ swaps) {
    __shared__ double s_val[kPanelThreads / 64];
    __shared__ int s_idx[kPanelThreads / 64];
    __shared__ int s_piv;
    __shared__ double rowP[kPW], rowC[kPW];
    __shared__ int s_loc[kPW], s_row[2 * kPW], s_src[2 * kPW];
    const int tid = threadIdx.x;
    const int R = N - k0;
    double v[kPanelRegRows][kPW];
#pragma unroll
    for (int q = 0; q < kPanelRegRows; ++q) {
        const int r = tid + q * kPanelThreads;
        const double2* src = reinterpret_cast<const double2*>(A + (size_t)(k0 + min(r, R - 1)) * N + k0);
#pragma unroll
        for (int j = 0; j < kPW / 2; ++j) {
            const double2 t = src[j];
            v[q][2 * j] = r < R ? t.x : 0.0;
            v[q][2 * j + 1] = r < R ? t.y : 0.0;
        }
    }
    // the column loop unrolled at compile time: v[q][cc] and v[q][j] index registers
    static_for(std::make_integer_sequence<int, kPW>{}, [&](auto ccv) {
        constexpr int cc = decltype(ccv)::value;
        double best = -1.0;
        int bi = R;
#pragma unroll
        for (int q = 0; q < kPanelRegRows; ++q) {
            const int r = tid + q * kPanelThreads;
            const double a = fabs(v[q][cc]);
            if (r >= cc && r < R && a > best) { best = a; bi = r; }
        }
        const int p = panel_pivot(best, bi, s_val, s_idx, &s_piv);
#pragma unroll
        for (int q = 0; q < kPanelRegRows; ++q) {
            const int r = tid + q * kPanelThreads;
            if (r == p) {
#pragma unroll
                for (int j = 0; j < kPW; ++j) rowP[j] = v[q][j];
            }
            if (r == cc && p != cc) {
#pragma unroll
                for (int j = 0; j < kPW; ++j) rowC[j] = v[q][j];
            }
        }
        if (tid == 0) {
            piv[k0 + cc] = k0 + p;
            s_loc[cc] = p;
        }
        __syncthreads();
        if (p != cc) {
#pragma unroll
            for (int q = 0; q < kPanelRegRows; ++q) {
                const int r = tid + q * kPanelThreads;
                if (r == cc) {
#pragma unroll
                    for (int j = 0; j < kPW; ++j) v[q][j] = rowP[j];
                } else if (r == p) {
#pragma unroll
                    for (int j = 0; j < kPW; ++j) v[q][j] = rowC[j];
                }
            }
        }
        const double d = rowP[cc];
        const bool skip = fabs(d) < kPivTiny;
        const double inv = 1.0 / d;
#pragma unroll
        for (int q = 0; q < kPanelRegRows; ++q) {
            const int r = tid + q * kPanelThreads;
            if (r > cc && r < R) {
                const double l = skip ? 0.0 : v[q][cc] * inv;
                v[q][cc] = l;
                if (!skip) {
#pragma unroll
                    for (int j = cc + 1; j < kPW; ++j) v[q][j] -= l * rowP[j];
                }
            }
        }
        // rowP / rowC are rewritten only after the next pivot's two barriers
    });
#pragma unroll
    for (int q = 0; q < kPanelRegRows; ++q) {
        const int r = tid + q * kPanelThreads;
        if (r < R) {
            double2* dst = reinterpret_cast<double2*>(A + (size_t)(k0 + r) * N + k0);
#pragma unroll
            for (int j = 0; j < kPW / 2; ++j) dst[j] = double2{v[q][2 * j], v[q][2 * j + 1]};
        }
    }
    if (tid == 0) {  // net permutation of the rows the swaps touched (at most 2 x 16)
        int* row = s_row;
        int* src = s_src;
        int m = 0;
        auto slot = [&](int x) {
            for (int i = 0; i < m; ++i)
                if (row[i] == x) return i;
            row[m] = x;
            src[m] = x;
            return m++;
        };
        for (int cc = 0; cc < kPW; ++cc) {
            const int a = slot(cc), b = slot(s_loc[cc]);
            const int t = src[a];
            src[a] = src[b];
            src[b] = t;
        }
        int cnt = 0;
        for (int i = 0; i < m; ++i)
            if (src[i] != row[i]) {
                swaps[1 + 2 * cnt] = k0 + row[i];
                swaps[2 + 2 * cnt] = k0 + src[i];
                ++cnt;
            }
        swaps[0] = cnt;
    }
}

// The same panel for more rows than the registers hold: in global memory,
// the whole row (all N columns, and the row's entry of sigma) swapped in
// place, so `swaps` gets no pairs.
__global__ void __launch_bounds__(kPanelThreads)
k_lu_panel_global(double* __restrict__ A, int N, int k0, int* __restrict__ piv, int* __restrict__ swaps,
                  int* __restrict__ sigma) {
    __shared__ double s_val[kPanelThreads / 64];
    __shared__ int s_idx[kPanelThreads / 64];
    __shared__ int s_piv;
    __shared__ double urow[kPW];
    const int tid = threadIdx.x;
    const int R = N - k0;
    double* P = A + (size_t)k0 * N + k0;  // P[r * N + j]: row k0 + r, column k0 + j
    for (int cc = 0; cc < kPW; ++cc) {
        double best = -1.0;
        int bi = R;
        for (int r = cc + tid; r < R; r += kPanelThreads) {  // increasing rows
            const double a = fabs(P[(size_t)r * N + cc]);
            if (a > best) { best = a; bi = r; }
        }
        const int p = panel_pivot(best, bi, s_val, s_idx, &s_piv);
        if (tid == 0) {
            piv[k0 + cc] = k0 + p;
            if (p != cc) {
                const int t = sigma[k0 + cc];
                sigma[k0 + cc] = sigma[k0 + p];
                sigma[k0 + p] = t;
            }
        }
        if (p != cc)
            for (int j = tid; j < N; j += kPanelThreads) {
                const double t = A[(size_t)(k0 + cc) * N + j];
                A[(size_t)(k0 + cc) * N + j] = A[(size_t)(k0 + p) * N + j];
                A[(size_t)(k0 + p) * N + j] = t;
            }
        __syncthreads();
        if (tid < kPW) urow[tid] = P[(size_t)cc * N + tid];
        __syncthreads();
        const double d = urow[cc];
        const bool skip = fabs(d) < kPivTiny;
        const double inv = 1.0 / d;
        for (int r = cc + 1 + tid; r < R; r += kPanelThreads) {
            double& a = P[(size_t)r * N + cc];
            a = skip ? 0.0 : a * inv;
        }
        __syncthreads();
        const int w = kPW - cc - 1;
        if (!skip && w > 0) {
            for (int e = tid; e < (R - cc - 1) * w; e += kPanelThreads) {
                const int r = cc + 1 + e / w, j = cc + 1 + e % w;
                P[(size_t)r * N + j] -= P[(size_t)r * N + cc] * urow[j];
            }
        }
        __syncthreads();
    }
    if (tid == 0) swaps[0] = 0;
}

// The panel's row permutation applied to the columns outside it (the L part
// of earlier panels and the trailing matrix) and to sigma (the rows' original
// indices: P I has row i = e_sigma[i]): one thread per column (the last one
// takes sigma), every source value loaded before any store.
__global__ void k_lu_laswp(double* __restrict__ A, int N, int k0, const int* __restrict__ swaps,
                           int* __restrict__ sigma) {
    const int cnt = swaps[0];
    if (cnt == 0) return;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > N - kPW) return;
    if (j == N - kPW) {
        int t[2 * kPW];
#pragma unroll
        for (int i = 0; i < 2 * kPW; ++i)
            if (i < cnt) t[i] = sigma[swaps[2 + 2 * i]];
#pragma unroll
        for (int i = 0; i < 2 * kPW; ++i)
            if (i < cnt) sigma[swaps[1 + 2 * i]] = t[i];
        return;
    }
    if (j >= k0) j += kPW;  // skip the panel's own columns
    double v[2 * kPW];
#pragma unroll
    for (int i = 0; i < 2 * kPW; ++i)
        if (i < cnt) v[i] = A[(size_t)swaps[2 + 2 * i] * N + j];
#pragma unroll
    for (int i = 0; i < 2 * kPW; ++i)
        if (i < cnt) A[(size_t)swaps[1 + 2 * i] * N + j] = v[i];
}

// X[b0 .. b0+NB, cols] = L_bb^-1 X[...] with L_bb the unit lower triangle of
// the diagonal block LU[b0 .. b0+NB, b0 .. b0+NB]; one thread per column,
// the column in registers.
template <int NB>
__global__ void __launch_bounds__(256) k_trsm_lower_unit(const double* __restrict__ LU, int ldlu, int b0,
                                                         double* __restrict__ X, int ldx, int col0, int ncols) {
    __shared__ double L[NB][NB];
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) L[e / NB][e % NB] = LU[(size_t)(b0 + e / NB) * ldlu + b0 + e % NB];
    __syncthreads();
    const int j = col0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= col0 + ncols) return;
    double x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = X[(size_t)(b0 + i) * ldx + j];
#pragma unroll
    for (int i = 1; i < NB; ++i) {
        double s = x[i];
#pragma unroll
        for (int m = 0; m < i; ++m) s -= L[i][m] * x[m];
        x[i] = s;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) X[(size_t)(b0 + i) * ldx + j] = x[i];
}

// X[b0 .. b0+64, cols] = U_bb^-1 X[...] by back substitution with the
// reference's rule x_i = 0 when |u_ii| <= 1e-14 (amg.cpp:285-290).
__global__ void __launch_bounds__(256) k_trsm_upper(const double* __restrict__ LU, int ldlu, int b0,
                                                   double* __restrict__ X, int ldx, int col0, int ncols) {
    __shared__ double U[kLuNB][kLuNB];
    for (int e = threadIdx.x; e < kLuNB * kLuNB; e += blockDim.x)
        U[e / kLuNB][e % kLuNB] = LU[(size_t)(b0 + e / kLuNB) * ldlu + b0 + e % kLuNB];
    __syncthreads();
    const int j = col0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= col0 + ncols) return;
    double x[kLuNB];
#pragma unroll
    for (int i = 0; i < kLuNB; ++i) x[i] = X[(size_t)(b0 + i) * ldx + j];
#pragma unroll
    for (int i = kLuNB - 1; i >= 0; --i) {
        double s = x[i];
#pragma unroll
        for (int m = i + 1; m < kLuNB; ++m) s -= U[i][m] * x[m];
        const double d = U[i][i];
        x[i] = fabs(d) > kPivTiny ? s / d : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kLuNB; ++i) X[(size_t)(b0 + i) * ldx + j] = x[i];
}

// C[m x n] -= A[m x k] B[k x n], row-major, m and n multiples of 16 (tiles
// past them load zeros and store nothing), k a multiple of 16.  A workgroup (4 waves) owns a 64 x 64 tile of C, each wave
// a 32 x 32 quarter as 2 x 2 MFMA tiles of 16 x 16; 16-deep slices of A and B
// are staged in LDS.  Operand layout of v_mfma_f64_16x16x4_f64: lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; register i of the result
// holds D[4 i + l / 16][l % 16].
__global__ void __launch_bounds__(256) k_gemm_sub_f64(int m, int n, int k, const double* __restrict__ A, int lda,
                                                      const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                      int ldc) {
    __shared__ double As[64][16 + 1];
    __shared__ double Bs[16][64 + 1];
    const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    v4d acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int kk = 0; kk < k; kk += 16) {
#pragma unroll
        for (int e = tid; e < 64 * 16; e += 256) {
            const int r = bm + (e >> 4);
            As[e >> 4][e & 15] = r < m ? A[(size_t)r * lda + kk + (e & 15)] : 0.0;
        }
#pragma unroll
        for (int e = tid; e < 16 * 64; e += 256) {
            const int c = bn + (e & 63);
            Bs[e >> 6][e & 63] = c < n ? B[(size_t)(kk + (e >> 6)) * ldb + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < 16; k4 += 4) {
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[wm + 16 * i + lr][k4 + lk];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[k4 + lk][wn + 16 * j + lr];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm + wm + 16 * i + 4 * r + lk, col = bn + wn + 16 * j + lr;
                if (row < m && col < n) C[(size_t)row * ldc + col] -= acc[i][j][r];
            }
}

// X[i][j] = (sigma[i] == j): X = P I.
__global__ void k_perm_identity(int N, const int* __restrict__ sigma, double* __restrict__ X) {
    const size_t total = (size_t)N * N;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / N), j = (int)(e % N);
        X[e] = sigma[i] == j ? 1.0 : 0.0;
    }
}

static void gemm_sub(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                     hipStream_t s) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_gemm_sub_f64, dim3((n + 63) / 64, (m + 63) / 64), dim3(256), 0, s, m, n, k, A, lda, B, ldb,
                       C, ldc);
    HSPS_CHECK_LAUNCH();
}

int dense_pad(int n) { return (std::max(n, 1) + kLuNB - 1) / kLuNB * kLuNB; }

void dense_inverse_device(Ctx* ctx, const Csr& A, DevBuf& G, int* ld_out, hipStream_t s) {
    const int n = A.rows, N = dense_pad(n);
    const size_t bytes = (size_t)N * N * sizeof(double);
    DevBuf LU(ctx, bytes), pivb(ctx, (size_t)N * sizeof(int)), sig(ctx, (size_t)N * sizeof(int)),
        swapb(ctx, (size_t)(1 + 4 * kPW) * sizeof(int));
    G = DevBuf(ctx, bytes);
    double* D = LU.as<double>();
    int* piv = pivb.as<int>();
    int* swaps = swapb.as<int>();
    int* sigma = sig.as<int>();
    const int gfill = (int)std::min<size_t>(((size_t)N * N + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_dense_init, dim3(gfill), dim3(kBlock), 0, s, N, n, D);
    hipLaunchKernelGGL(k_dense_scatter, dim3((n + 3) / 4), dim3(kBlock), 0, s, n, N, A.rp.as<int>(), A.ci.as<int>(),
                       A.v64.as<double>(), D);
    hipLaunchKernelGGL(k_iota, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, N, sigma);
    HSPS_CHECK_LAUNCH();
    // P A = L U, 16 columns at a time
    for (int k0 = 0; k0 < N; k0 += kPW) {
        const int R = N - k0;
        if (R <= kPanelRegRows * kPanelThreads)
            hipLaunchKernelGGL(k_lu_panel_reg, dim3(1), dim3(kPanelThreads), 0, s, D, N, k0, piv, swaps);
        else
            hipLaunchKernelGGL(k_lu_panel_global, dim3(1), dim3(kPanelThreads), 0, s, D, N, k0, piv, swaps, sigma);
        HSPS_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_lu_laswp, dim3((N - kPW + 1 + 255) / 256), dim3(256), 0, s, D, N, k0, (const int*)swaps,
                           sigma);
        HSPS_CHECK_LAUNCH();
        const int rest = R - kPW;
        if (rest > 0) {
            hipLaunchKernelGGL(k_trsm_lower_unit<kPW>, dim3((rest + 255) / 256), dim3(256), 0, s, (const double*)D, N,
                               k0, D, N, k0 + kPW, rest);
            HSPS_CHECK_LAUNCH();
            // A22 -= L21 U12
            gemm_sub(rest, rest, kPW, D + (size_t)(k0 + kPW) * N + k0, N, D + (size_t)k0 * N + k0 + kPW, N,
                     D + (size_t)(k0 + kPW) * N + k0 + kPW, N, s);
        }
    }
    // G = U^-1 L^-1 P
    double* X = G.as<double>();
    hipLaunchKernelGGL(k_perm_identity, dim3(gfill), dim3(kBlock), 0, s, N, (const int*)sigma, X);
    HSPS_CHECK_LAUNCH();
    const int nb = N / kLuNB;
    for (int b = 0; b < nb; ++b) {  // forward: L Y = P I
        const int b0 = b * kLuNB;
        gemm_sub(kLuNB, N, b0, D + (size_t)b0 * N, N, X, N, X + (size_t)b0 * N, N, s);
        hipLaunchKernelGGL(k_trsm_lower_unit<kLuNB>, dim3((N + 255) / 256), dim3(256), 0, s, (const double*)D, N, b0, X, N, 0,
                           N);
        HSPS_CHECK_LAUNCH();
    }
    for (int b = nb - 1; b >= 0; --b) {  // backward: U G = Y
        const int b0 = b * kLuNB, e0 = b0 + kLuNB;
        gemm_sub(kLuNB, N, N - e0, D + (size_t)b0 * N + e0, N, X + (size_t)e0 * N, N, X + (size_t)b0 * N, N, s);
        hipLaunchKernelGGL(k_trsm_upper, dim3((N + 255) / 256), dim3(256), 0, s, (const double*)D, N, b0, X, N, 0, N);
        HSPS_CHECK_LAUNCH();
    }
    HSPS_HIP(hipStreamSynchronize(s));  // LU / pivots / sigma are freed on return
    *ld_out = N;
}

}  // namespace hsps_dev
