// =============================================================================
// krylov.hip — device-resident PCG and FGMRES(m).
//
// PCG (cg_solver.cpp:37-191), per iteration:
//   K1  q = A p fused with the p.q partial sums      F1  alpha (or p.q breakdown, :109-120)
//   K2  x += alpha p, r -= alpha q, z = D^-1 r with the r.r / r.z partials
//                                                    F2  rel = |r|/|b|, rho, beta and the
//                                                        convergence / rho-breakdown /
//                                                        iteration-limit decisions
//   K3  p = z + beta p
// F1/F2 are one-workgroup finish kernels that sum the partials in a fixed
// order (deterministic) and update the device scalar state.
// FGMRES(m) (fgmres_solver.cpp:36-241): right-preconditioned, modified
// Gram-Schmidt with each axpy fused into the next dot, Givens rotations and
// the least-squares residual kept on the device in FP64.
//
// Precision: HSPS_FP64 runs everything in FP64.  HSPS_MIXED runs the SpMV,
// preconditioner and Krylov vectors in FP32 with FP64 reductions and FP64
// H/Givens; the FP64 true residual b - A x is recomputed
//   * CG: by reliable updates — whenever |r| falls below delta * max|r|
//     since the last one, and to confirm convergence;
//   * FGMRES: at every restart (each cycle is one refinement step), and a
//     solve only ends when a cycle starts below tol.
// The host polls the device state every check_every iterations, one chunk
// ahead, so the GPU never waits on the host within a chunk.
// =============================================================================
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <limits>
#include <type_traits>

#include "device_common.hpp"
#include "precond.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

enum : int { FLAG_NONE = 0, FLAG_BREAK_PQ = 1, FLAG_BREAK_RHO = 2, FLAG_RELIABLE = 3, FLAG_CHECK = 4 };

struct CgState {
    double rho, rho_old, alpha, beta, rel, bnorm, tol, rmax, delta;
    double rr0, bb, rr_true;
    int iter, max_iter, halt, done, flag, restarts, reliable, pad;
};

struct GmState {
    double beta2, beta, res, tol_abs, inv_h;
    double inner_rel;  // mixed: a cycle also ends once res < inner_rel * beta (0 in FP64)
    int jdone, cycle_done, m, pad;
};

// ---- CG kernels -----------------------------------------------------------------------
template <class T>
struct EpiCgPq {
    static constexpr int NV = 1;
    const T* p;
    T* q;
    const CgState* st;
    double* partials;
    __device__ bool skip() const { return st->halt != 0; }
    template <class A> __device__ void row(int i, A s, double* red) const {
        const T qi = (T)s;
        q[i] = qi;
        red[0] += (double)p[i] * (double)qi;
    }
};

struct OpAlpha {  // F1
    CgState* st;
    __device__ bool skip() const { return st->halt != 0; }
    __device__ void operator()(const double* tot) const {
        const double pq = tot[0];
        if (fabs(pq) < kEps * 1e6 * fabs(st->rho)) {  // cg_solver.cpp:109
            st->flag = FLAG_BREAK_PQ;
            st->iter += 1;
            st->halt = 1;
        } else {
            st->alpha = st->rho / pq;
        }
    }
};

// PREC: 0 none (z = r), 1 Jacobi (z = dinv r, inline), 2 external (z computed
// after this kernel).  XT: type of the x being updated (FP64 x, or the FP32
// low-order accumulator in mixed mode).
template <class T, class XT, int PREC>
__global__ void __launch_bounds__(kBlock)
k_cg_update(int n, XT* __restrict__ x, const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ r,
            const T* __restrict__ dinv, const CgState* st, double* partials) {
    if (st->halt) return;
    const double alpha = st->alpha;
    double red[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = (XT)((double)x[i] + alpha * (double)p[i]);
        const T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        const double rd = (double)ri;
        red[0] += rd * rd;
        if constexpr (PREC == 1) red[1] += rd * (double)(T)(dinv[i] * ri);
        if constexpr (PREC == 0) red[1] += rd * rd;
    }
    block_partial<2>(red, partials);
}

template <int PREC, bool MIXED>
struct OpUpdate {  // F2
    CgState* st;
    __device__ bool skip() const { return st->halt != 0; }
    __device__ void operator()(const double* tot) const {
        const double rn = sqrt(tot[0]);
        const double rel = rn / st->bnorm;
        st->rel = rel;
        st->iter += 1;
        if (rel < st->tol) {  // cg_solver.cpp:143
            if (MIXED) st->flag = FLAG_CHECK; else st->done = 1;
            st->halt = 1;
            return;
        }
        if (MIXED) {
            st->rmax = fmax(st->rmax, rn);
            if (rn < st->delta * st->rmax) {
                st->flag = FLAG_RELIABLE;
                st->halt = 1;
                return;
            }
        }
        if (PREC == 2) return;  // rho from the r.z pass after the external apply
        st->rho_old = st->rho;
        st->rho = tot[1];
        if (fabs(st->rho_old) < kEps * 1e8 * fabs(st->rho)) {  // cg_solver.cpp:157
            st->flag = FLAG_BREAK_RHO;
            st->halt = 1;
            return;
        }
        st->beta = st->rho / st->rho_old;
        if (st->iter >= st->max_iter) st->halt = 1;
    }
};

template <class T>
__global__ void __launch_bounds__(kBlock)
k_dot_t(int n, const T* __restrict__ a, const T* __restrict__ b, const int* skip, double* partials) {
    if (skip && *skip) return;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) red[0] += (double)a[i] * (double)b[i];
    block_partial<1>(red, partials);
}

struct OpRz {  // rho = r.z for an external preconditioner
    CgState* st;
    __device__ bool skip() const { return st->halt != 0; }
    __device__ void operator()(const double* tot) const {
        st->rho_old = st->rho;
        st->rho = tot[0];
        if (fabs(st->rho_old) < kEps * 1e8 * fabs(st->rho)) {
            st->flag = FLAG_BREAK_RHO;
            st->halt = 1;
            return;
        }
        st->beta = st->rho / st->rho_old;
        if (st->iter >= st->max_iter) st->halt = 1;
    }
};

// p = z + beta p (cg_solver.cpp:171); with INIT: p = z and the r.z partials
// for rho (cg_solver.cpp:77-85) from a freshly formed r.
template <class T, int PREC, bool INIT>
__global__ void __launch_bounds__(kBlock)
k_cg_p(int n, T* __restrict__ p, const T* __restrict__ r, const T* __restrict__ z, const T* __restrict__ dinv,
       const CgState* st, double* partials) {
    if (!INIT && st->halt) return;
    const double beta = INIT ? 0.0 : st->beta;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        T zi;
        if constexpr (PREC == 0) zi = r[i];
        else if constexpr (PREC == 1) zi = (T)(dinv[i] * r[i]);
        else zi = z[i];
        if (INIT) {
            p[i] = zi;
            red[0] += (double)r[i] * (double)zi;
        } else {
            p[i] = (T)((double)zi + beta * (double)p[i]);
        }
    }
    if constexpr (INIT) block_partial<1>(red, partials);
}

struct OpInitRho {
    CgState* st;
    __device__ bool skip() const { return false; }
    __device__ void operator()(const double* tot) const {
        st->rho = tot[0];
        st->rho_old = tot[0];
    }
};

// Mixed: x64 += x_lo; x_lo = 0.
__global__ void k_flush(int n, double* __restrict__ x, float* __restrict__ xlo) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += (double)xlo[i];
        xlo[i] = 0.0f;
    }
}

// Mixed: replace r by the FP64 true residual rounded to FP32 (r.z partials).
template <int PREC>
__global__ void __launch_bounds__(kBlock)
k_cg_reliable(int n, const double* __restrict__ r64, float* __restrict__ r, const float* __restrict__ dinv,
              double* partials) {
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float ri = (float)r64[i];
        r[i] = ri;
        if constexpr (PREC == 1) red[0] += (double)ri * (double)(dinv[i] * ri);
        if constexpr (PREC == 0) red[0] += (double)ri * (double)ri;
    }
    block_partial<1>(red, partials);
}

// Rebuild rho / beta from the replaced r; the p direction is kept
// (reliable update).  A CHECK whose true residual is below tol ends the solve.
template <int PREC>
struct OpReliable {
    CgState* st;
    __device__ bool skip() const { return false; }
    __device__ void operator()(const double* tot) const {
        const double rn = sqrt(st->rr_true);
        const double rel = rn / st->bnorm;
        st->reliable += 1;
        st->rmax = rn;
        if (st->flag == FLAG_CHECK && rel < st->tol) {
            st->rel = rel;
            st->done = 1;
            st->halt = 1;
            st->flag = FLAG_NONE;
            return;
        }
        st->flag = FLAG_NONE;
        st->halt = 0;
        if (PREC != 2) {
            st->rho_old = st->rho;
            st->rho = tot[0];
            st->beta = st->rho / st->rho_old;
            if (st->iter >= st->max_iter) st->halt = 1;
        }
    }
};

// ---- FGMRES kernels ---------------------------------------------------------------------
// V0 = r / beta, or stop when beta < tol_abs (fgmres_solver.cpp:82-96).
template <class T>
__global__ void k_gm_start(int n, const double* __restrict__ r, T* __restrict__ v0, GmState* st, double* g) {
    const double beta = sqrt(st->beta2);
    const bool stop = beta < st->tol_abs;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->beta = beta;
        st->res = beta;
        st->jdone = 0;
        st->cycle_done = stop ? 1 : 0;
        for (int k = 0; k <= st->m; ++k) g[k] = 0.0;
        g[0] = beta;
    }
    if (stop) return;
    const double s = 1.0 / beta;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v0[i] = (T)(r[i] * s);
}

// One modified Gram-Schmidt step, with the previous step's axpy fused in:
//   w -= h_prev * v_prev  (when h_prev given);  partials of w . v_i
template <class T>
__global__ void __launch_bounds__(kBlock)
k_mgs(int n, T* __restrict__ w, const T* __restrict__ vprev, const double* hprev, const T* __restrict__ vi,
      const int* skip, double* partials) {
    if (*skip) return;
    const double h = hprev ? *hprev : 0.0;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        T wi = w[i];
        if (hprev) {
            wi = (T)((double)wi - h * (double)vprev[i]);
            w[i] = wi;
        }
        red[0] += (double)wi * (double)vi[i];
    }
    block_partial<1>(red, partials);
}

struct OpStoreH {  // H[i][j] = w . v_i
    double* h;
    const int* flag;
    __device__ bool skip() const { return *flag != 0; }
    __device__ void operator()(const double* tot) const { *h = tot[0]; }
};

// w -= H[j][j] v_j ; partials of |w|^2
template <class T>
__global__ void __launch_bounds__(kBlock)
k_mgs_last(int n, const double* hjj, T* __restrict__ w, const T* __restrict__ vj, const GmState* st,
           double* partials) {
    if (st->cycle_done) return;
    const double h = *hjj;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const T wi = (T)((double)w[i] - h * (double)vj[i]);
        w[i] = wi;
        red[0] += (double)wi * (double)wi;
    }
    block_partial<1>(red, partials);
}

__device__ void gm_givens(double f, double g, double& cs, double& sn, double& r) {  // fgmres_solver.cpp:36-45
    if (fabs(g) < kEps) { cs = 1.0; sn = 0.0; r = f; }
    else if (fabs(f) < kEps) { cs = 0.0; sn = 1.0; r = g; }
    else { const double t = sqrt(f * f + g * g); cs = f / t; sn = g / t; r = t; }
}
__device__ void gm_rot(double& dx, double& dy, double cs, double sn) {  // fgmres_solver.cpp:47-51
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

// H[j+1][j] = |w|; Givens update of column j and the least-squares residual
// estimate (fgmres_solver.cpp:126-148).
struct OpGivens {
    int j;
    double* H;
    double* cs;
    double* sn;
    double* g;
    GmState* st;
    __device__ bool skip() const { return st->cycle_done != 0; }
    __device__ void operator()(const double* tot) const {
        const int m = st->m;
        const double hn = sqrt(tot[0]);
        H[(j + 1) * m + j] = hn;
        st->inv_h = hn > kEps ? 1.0 / hn : 0.0;  // lucky breakdown leaves V_{j+1} = 0
        for (int i = 0; i < j; ++i) gm_rot(H[i * m + j], H[(i + 1) * m + j], cs[i], sn[i]);
        double rv;
        gm_givens(H[j * m + j], H[(j + 1) * m + j], cs[j], sn[j], rv);
        H[j * m + j] = rv;
        H[(j + 1) * m + j] = 0.0;
        gm_rot(g[j], g[j + 1], cs[j], sn[j]);
        st->res = fabs(g[j + 1]);
        st->jdone = j + 1;
        if (st->res < st->tol_abs || st->res < st->inner_rel * st->beta) st->cycle_done = 1;
    }
};

// ---- pipelined FGMRES: classical Gram-Schmidt with one re-orthogonalisation
// (CGS2), each pass two launches whatever j is: the dots of w against all of
// V_0..V_j at once (blockIdx.y = i), then w -= sum_i h_i V_i in index order
// (the second pass also returns |w|^2 partials).  H(:, j) = h1 + h2.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_cgs_dots(int n, const T* __restrict__ w, const T* __restrict__ V, size_t ld, const int* skip, double* partials) {
    if (*skip) return;
    const T* vi = V + (size_t)blockIdx.y * ld;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) red[0] += (double)w[i] * (double)vi[i];
    block_partial<1>(red, partials + (size_t)blockIdx.y * gridDim.x);
}

// out[i * stride] = sum of partials[i * g .. i * g + g) in a fixed order, one
// wave per i.
__global__ void __launch_bounds__(kBlock)
k_cgs_finish(int nv, int g, const double* __restrict__ partials, double* out, int stride, const int* skip) {
    if (*skip) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i < nv; i += kBlock / 64) {
        double a = 0.0;
        for (int b = lane; b < g; b += 64) a += partials[(size_t)i * g + b];
        a = wave_sum(a);
        if (lane == 0) out[(size_t)i * stride] = a;
    }
}

template <class T, bool NORM>
__global__ void __launch_bounds__(kBlock)
k_cgs_update(int n, T* __restrict__ w, const T* __restrict__ V, size_t ld, int nv, const double* h, int stride,
             const int* skip, double* partials) {
    if (*skip) return;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double a = (double)w[i];
        for (int c = 0; c < nv; ++c) a -= h[(size_t)c * stride] * (double)V[(size_t)c * ld + i];
        const T wi = (T)a;
        w[i] = wi;
        if (NORM) red[0] += (double)wi * (double)wi;
    }
    if (NORM) block_partial<1>(red, partials);
}

// H(:, j) += h2; then as OpGivens, except that a breakdown (|w| <= eps) ends
// the cycle before column j is rotated (pipelined_fgmres_solver.cpp:128-136).
struct OpGivensPipe {
    OpGivens base;
    const double* h2;
    __device__ bool skip() const { return base.skip(); }
    __device__ void operator()(const double* tot) const {
        const int j = base.j, m = base.st->m;
        for (int i = 0; i <= j; ++i) base.H[i * m + j] += h2[i];
        const double hn = sqrt(tot[0]);
        if (!(hn > kEps)) {
            base.H[(j + 1) * m + j] = hn;
            base.st->inv_h = 0.0;
            base.st->jdone = j + 1;
            base.st->cycle_done = 1;
            return;
        }
        base(tot);
    }
};

template <class T>
__global__ void k_gm_scale(int n, T* __restrict__ v, const GmState* st) {
    if (st->cycle_done) return;
    const double s = st->inv_h;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = (T)((double)v[i] * s);
}

// y = H(0:k,0:k)^-1 g (fgmres_solver.cpp:57-66), one thread.
__global__ void k_gm_backsolve(const double* H, const double* g, double* y, const GmState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int k = st->jdone, m = st->m;
    for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int c = i + 1; c < k; ++c) s -= H[i * m + c] * y[c];
        y[i] = fabs(H[i * m + i]) > kEps ? s / H[i * m + i] : 0.0;
    }
}

// x += sum_i y_i Z_i in the reference's order (fgmres_solver.cpp:156-157).
template <class T>
__global__ void k_gm_xupdate(int n, double* __restrict__ x, const T* __restrict__ Z, size_t ld, const double* y,
                             const GmState* st) {
    const int k = st->jdone;
    if (k == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = x[i];
        for (int c = 0; c < k; ++c) a += y[c] * (double)Z[c * ld + i];
        x[i] = a;
    }
}

// ---- host side -------------------------------------------------------------------------
struct Pinned {
    void* p = nullptr;
    explicit Pinned(size_t b) { HSPS_HIP(hipHostMalloc(&p, b, hipHostMallocDefault)); }
    ~Pinned() { if (p) (void)hipHostFree(p); }
};
struct Event {
    hipEvent_t e = nullptr;
    Event() { HSPS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    ~Event() { if (e) (void)hipEventDestroy(e); }
};

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void finish_stats(hsps_stats* s, long long F, long long B) {
    s->flop_count = F;
    s->mem_bytes = B;
    s->energy_joules = 2.0e-10 * (double)F + 5.0e-9 * (double)B;
}

// True relative residual ||b - A x|| / ||b|| in FP64 (device), for stats.
static double true_rel(const Csr& A, const double* b, const double* x, double* r64, double bnorm, double* d_out,
                       double* partials, hipStream_t s) {
    residual_f64(A, b, x, r64, s, d_out, partials, nullptr);
    double h = 0.0;
    HSPS_HIP(hipMemcpyAsync(&h, d_out, sizeof(double), hipMemcpyDeviceToHost, s));
    HSPS_HIP(hipStreamSynchronize(s));
    return bnorm > 0 ? std::sqrt(h) / bnorm : std::sqrt(h);
}

// Buffers of one PCG solve: device workspace, two host slots for the state
// and, optionally, two events (without them the host polls synchronously).
struct CgWork {
    void *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;  // working precision, n each
    double* r64 = nullptr;       // FP64 residual, n
    float* xlo = nullptr;        // mixed: FP32 low-order part of x, n
    double* partials = nullptr;  // 2 x max(reduction blocks, SpMV row blocks)
    CgState* st = nullptr;
    double* scal = nullptr;      // 8 doubles
    CgState* host = nullptr;     // 2 host slots
    hipEvent_t ev[2] = {nullptr, nullptr};
};
struct WorkView {
    void* p;
    template <class U> U* as() const { return static_cast<U*>(p); }
};
// 0 none, 1 Jacobi (inline), 2 any other preconditioner (z buffer + apply)
static int cg_prec(const Precond* M) {
    const int kind = M ? M->kind : HSPS_PRECOND_NONE;
    return kind == HSPS_PRECOND_NONE ? 0 : kind == HSPS_PRECOND_JACOBI ? 1 : 2;
}
// Device bytes of a PCG solve's workspace; with base != null, also points
// w's device buffers into [base, base + bytes), each 256-byte aligned.
static size_t cg_work_layout(const Csr& A, bool mixed, int prec, unsigned char* base, CgWork* w) {
    const size_t nb = (size_t)std::max(A.rows, 1);
    const size_t vs = mixed ? sizeof(float) : sizeof(double);
    const size_t np = 2 * (size_t)std::max(std::max(reduction_blocks(A.rows), A.n_blocks), 1);
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* ptr = base && bytes ? base + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return ptr;
    };
    void* r = take(nb * vs);
    void* p = take(nb * vs);
    void* q = take(nb * vs);
    void* z = take(prec == 2 ? nb * vs : 0);
    void* r64 = take(nb * sizeof(double));
    void* xlo = take(mixed ? nb * sizeof(float) : 0);
    void* part = take(np * sizeof(double));
    void* st = take(sizeof(CgState));
    void* scal = take(8 * sizeof(double));
    if (base) {
        w->r = r; w->p = p; w->q = q; w->z = z;
        w->r64 = static_cast<double*>(r64);
        w->xlo = static_cast<float*>(xlo);
        w->partials = static_cast<double*>(part);
        w->st = static_cast<CgState*>(st);
        w->scal = static_cast<double*>(scal);
    }
    return off;
}

template <class T>
static int run_cg(const Csr& A, const Precond* M, const double* b, double* x, const hsps_params& p, hsps_stats* out,
                  const CgWork& w) {
    constexpr bool MIXED = std::is_same<T, float>::value;
    Ctx* ctx = A.ctx;
    hipStream_t s = ctx->stream;
    const int n = A.rows;
    const double t0 = now_s();
    hsps_stats st{};
    const int PREC = cg_prec(M);
    if (MIXED && !A.v32.p && A.nnz) throw Error(HSPS_ERR_UNSUPPORTED, "matrix has no FP32 copy");

    const size_t nb = (size_t)std::max(n, 1);
    const WorkView r{w.r}, pv{w.p}, q{w.q}, z{w.z}, r64{w.r64}, xlo{w.xlo}, scal{w.scal};
    const int g = reduction_blocks(n);
    CgState* hst = w.host;
    CgState* d = w.st;
    double* part = w.partials;

    // r = b - A x, ||r||^2, ||b||^2  (cg_solver.cpp:52-57)
    residual_f64(A, b, x, r64.as<double>(), s, &d->rr0, part, nullptr);
    dot_f64_dev(n, b, b, &d->bb, part, s);
    HSPS_HIP(hipMemcpyAsync(hst, d, sizeof(CgState), hipMemcpyDeviceToHost, s));
    HSPS_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(hst->bb), r0 = std::sqrt(hst->rr0);
    st.initial_residual = r0;
    auto finish = [&](int conv) {
        st.converged = conv;
        st.solve_time_s = now_s() - t0;
        if (out) *out = st;
        return conv;
    };
    if (bnorm < kEps) {  // cg_solver.cpp:60-66 (absolute residual reported)
        st.final_residual = r0;
        st.true_residual = r0;
        return finish(1);
    }
    if (r0 / bnorm < p.tol) {
        st.final_residual = st.true_residual = r0 / bnorm;
        return finish(1);
    }

    CgState init{};
    init.bnorm = bnorm;
    init.tol = p.tol;
    init.max_iter = p.max_iter;
    init.rmax = r0;
    init.delta = p.reliable_delta > 0 ? p.reliable_delta : 0.1;
    init.rr0 = hst->rr0;
    init.bb = hst->bb;
    init.halt = p.max_iter <= 0 ? 1 : 0;
    std::memcpy(hst, &init, sizeof init);
    HSPS_HIP(hipMemcpyAsync(d, hst, sizeof(CgState), hipMemcpyHostToDevice, s));

    const T* dinv = nullptr;
    if (PREC == 1) dinv = MIXED ? (const T*)precond_jacobi_dinv32(*M) : (const T*)precond_jacobi_dinv64(*M);
    T* rv = r.as<T>();
    T* p_ = pv.as<T>();
    T* q_ = q.as<T>();
    T* z_ = z.as<T>();

    // r (working precision), z = M r, p = z, rho = r.z  (cg_solver.cpp:77-85)
    auto start_direction = [&]() {
        if (MIXED) copy_f64_to_f32(n, r64.as<double>(), (float*)rv, s);
        else HSPS_HIP(hipMemcpyAsync(rv, r64.p, nb * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (PREC == 2) {
            if (MIXED) precond_apply_f32(*M, (const float*)rv, (float*)z_, s, nullptr);
            else precond_apply_f64(*M, (const double*)rv, (double*)z_, s, nullptr);
        }
        switch (PREC) {
            case 0: hipLaunchKernelGGL((k_cg_p<T, 0, true>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
            case 1: hipLaunchKernelGGL((k_cg_p<T, 1, true>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
            default: hipLaunchKernelGGL((k_cg_p<T, 2, true>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
        }
        launch_finish<1>(part, g, OpInitRho{d}, s);
        HSPS_CHECK_LAUNCH();
    };
    start_direction();
    if (MIXED) fill_zero_f32(n, xlo.as<float>(), s);

    auto spmv_pq = [&]() {
        EpiCgPq<T> e;
        e.p = p_; e.q = q_; e.st = d; e.partials = part;
        if (MIXED) launch_spmv(A, A.v32.as<float>(), (const float*)p_, e, s);
        else launch_spmv(A, A.v64.as<double>(), (const double*)p_, e, s);
        launch_finish<1>(part, spmv_blocks<T>(A), OpAlpha{d}, s);
    };
    auto update = [&]() {
        if (MIXED) {
            float* xl = xlo.as<float>();
            switch (PREC) {
                case 0: hipLaunchKernelGGL((k_cg_update<T, float, 0>), dim3(g), dim3(kBlock), 0, s, n, xl, p_, q_, rv, dinv, d, part); break;
                case 1: hipLaunchKernelGGL((k_cg_update<T, float, 1>), dim3(g), dim3(kBlock), 0, s, n, xl, p_, q_, rv, dinv, d, part); break;
                default: hipLaunchKernelGGL((k_cg_update<T, float, 2>), dim3(g), dim3(kBlock), 0, s, n, xl, p_, q_, rv, dinv, d, part); break;
            }
        } else {
            switch (PREC) {
                case 0: hipLaunchKernelGGL((k_cg_update<T, double, 0>), dim3(g), dim3(kBlock), 0, s, n, x, p_, q_, rv, dinv, d, part); break;
                case 1: hipLaunchKernelGGL((k_cg_update<T, double, 1>), dim3(g), dim3(kBlock), 0, s, n, x, p_, q_, rv, dinv, d, part); break;
                default: hipLaunchKernelGGL((k_cg_update<T, double, 2>), dim3(g), dim3(kBlock), 0, s, n, x, p_, q_, rv, dinv, d, part); break;
            }
        }
        switch (PREC) {
            case 0: launch_finish<2>(part, g, OpUpdate<0, MIXED>{d}, s); break;
            case 1: launch_finish<2>(part, g, OpUpdate<1, MIXED>{d}, s); break;
            default: launch_finish<2>(part, g, OpUpdate<2, MIXED>{d}, s); break;
        }
        HSPS_CHECK_LAUNCH();
        if (PREC == 2) {
            if (MIXED) precond_apply_f32(*M, (const float*)rv, (float*)z_, s, &d->halt);
            else precond_apply_f64(*M, (const double*)rv, (double*)z_, s, &d->halt);
            hipLaunchKernelGGL(k_dot_t<T>, dim3(g), dim3(kBlock), 0, s, n, (const T*)rv, (const T*)z_, &d->halt, part);
            launch_finish<1>(part, g, OpRz{d}, s);
            HSPS_CHECK_LAUNCH();
        }
    };
    auto pupdate = [&]() {
        switch (PREC) {
            case 0: hipLaunchKernelGGL((k_cg_p<T, 0, false>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
            case 1: hipLaunchKernelGGL((k_cg_p<T, 1, false>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
            default: hipLaunchKernelGGL((k_cg_p<T, 2, false>), dim3(g), dim3(kBlock), 0, s, n, p_, rv, z_, dinv, d, part); break;
        }
        HSPS_CHECK_LAUNCH();
    };

    const int chunk = std::max(1, p.check_every > 0 ? p.check_every : 32);
    int k = 0;
    bool pending = false;
    long long pq_breaks = 0, rho_breaks = 0;
    auto read_state = [&]() {
        HSPS_HIP(hipMemcpyAsync(hst, d, sizeof(CgState), hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
    };
    for (;;) {
        for (int c = 0; c < chunk; ++c) {
            spmv_pq();
            update();
            pupdate();
        }
        CgState* slot = hst + (k & 1);
        HSPS_HIP(hipMemcpyAsync(slot, d, sizeof(CgState), hipMemcpyDeviceToHost, s));
        bool halted = false;
        if (!w.ev[0]) {  // no events: read this chunk's state at once
            HSPS_HIP(hipStreamSynchronize(s));
            halted = slot->halt != 0;
        } else {  // read the previous chunk's state while this one runs
            HSPS_HIP(hipEventRecord(w.ev[k & 1], s));
            if (pending) {
                HSPS_HIP(hipEventSynchronize(w.ev[(k - 1) & 1]));
                halted = hst[(k - 1) & 1].halt != 0;
            }
        }
        pending = true;
        ++k;
        if (!halted) continue;
        read_state();  // everything after the halt was a no-op
        pending = false;
        k = 0;
        CgState cur = hst[0];
        const bool refine = MIXED && (cur.flag == FLAG_CHECK || cur.flag == FLAG_RELIABLE);
        if (cur.done || (cur.iter >= cur.max_iter && !refine)) break;
        if (cur.flag == FLAG_BREAK_PQ || cur.flag == FLAG_BREAK_RHO) {
            // restart from the current x (cg_solver.cpp:109-120, 157-167)
            if (cur.flag == FLAG_BREAK_PQ) ++pq_breaks; else ++rho_breaks;
            if (MIXED) hipLaunchKernelGGL(k_flush, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, xlo.as<float>());
            residual_f64(A, b, x, r64.as<double>(), s, nullptr, nullptr, nullptr);
            start_direction();
            cur.restarts += 1;
            cur.flag = FLAG_NONE;
            cur.halt = cur.iter >= cur.max_iter ? 1 : 0;
            int upd[3] = {cur.halt, 0, cur.flag};
            HSPS_HIP(hipMemcpyAsync(&d->halt, &upd[0], sizeof(int), hipMemcpyHostToDevice, s));
            HSPS_HIP(hipMemcpyAsync(&d->flag, &upd[2], sizeof(int), hipMemcpyHostToDevice, s));
            HSPS_HIP(hipMemcpyAsync(&d->restarts, &cur.restarts, sizeof(int), hipMemcpyHostToDevice, s));
            HSPS_HIP(hipStreamSynchronize(s));
            if (cur.halt) break;
            continue;
        }
        if (refine) {
            hipLaunchKernelGGL(k_flush, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, xlo.as<float>());
            residual_f64(A, b, x, r64.as<double>(), s, &d->rr_true, part, nullptr);
            switch (PREC) {
                case 0: hipLaunchKernelGGL(k_cg_reliable<0>, dim3(g), dim3(kBlock), 0, s, n, r64.as<double>(), (float*)rv, (const float*)dinv, part);
                        launch_finish<1>(part, g, OpReliable<0>{d}, s); break;
                case 1: hipLaunchKernelGGL(k_cg_reliable<1>, dim3(g), dim3(kBlock), 0, s, n, r64.as<double>(), (float*)rv, (const float*)dinv, part);
                        launch_finish<1>(part, g, OpReliable<1>{d}, s); break;
                default: hipLaunchKernelGGL(k_cg_reliable<2>, dim3(g), dim3(kBlock), 0, s, n, r64.as<double>(), (float*)rv, (const float*)dinv, part);
                        launch_finish<1>(part, g, OpReliable<2>{d}, s); break;
            }
            HSPS_CHECK_LAUNCH();
            if (PREC == 2) {
                precond_apply_f32(*M, (const float*)rv, (float*)z_, s, &d->halt);
                hipLaunchKernelGGL(k_dot_t<T>, dim3(g), dim3(kBlock), 0, s, n, (const T*)rv, (const T*)z_, &d->halt, part);
                launch_finish<1>(part, g, OpRz{d}, s);
                HSPS_CHECK_LAUNCH();
            }
            pupdate();
            read_state();
            if (hst[0].done || hst[0].halt) break;
            continue;
        }
        break;  // halted for the iteration limit
    }
    if (MIXED) hipLaunchKernelGGL(k_flush, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, xlo.as<float>());
    read_state();
    const CgState fin = hst[0];
    st.iterations = fin.iter;
    st.restarts = fin.restarts;
    st.reliable_updates = fin.reliable;
    st.final_residual = fin.rel;
    st.true_residual = true_rel(A, b, x, r64.as<double>(), bnorm, scal.as<double>(), part, s);
    const int conv = MIXED ? (st.true_residual < p.tol && fin.done) : (fin.rel < p.tol);
    const CgCost cc = cg_cost(n, A.nnz, fin.iter, fin.done != 0, pq_breaks, rho_breaks, MIXED ? 4 : 8);
    finish_stats(&st, cc.flops, cc.bytes);
    return finish(conv);
}

// PCG with its workspace allocated for the call (one device buffer, pinned
// host slots, two events).
template <class T>
static int run_cg_alloc(const Csr& A, const Precond* M, const double* b, double* x, const hsps_params& p,
                        hsps_stats* out) {
    constexpr bool MIXED = std::is_same<T, float>::value;
    const int prec = cg_prec(M);
    DevBuf ws(A.ctx, cg_work_layout(A, MIXED, prec, nullptr, nullptr));
    CgWork w;
    cg_work_layout(A, MIXED, prec, ws.as<unsigned char>(), &w);
    Pinned pin(sizeof(CgState) * 2);
    w.host = static_cast<CgState*>(pin.p);
    Event ev[2];
    w.ev[0] = ev[0].e;
    w.ev[1] = ev[1].e;
    return run_cg<T>(A, M, b, x, p, out, w);
}

// ---- PCG without allocation ---------------------------------------------------------
// The workspace of hsps_cg_solve_noalloc: one array per device, reserved when
// the code object loads and held until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns on each device (the
// symbol's name in host code is not a device address).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The scratch address on `device` (current), looked up on the device's first call.
static unsigned char* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(HSPS_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            throw Error(HSPS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    return g_scratch_ptr[device];
}

template <class T>
static int run_cg_noalloc(const Csr& A, const Precond* M, const double* b, double* x, const hsps_params& p,
                          hsps_stats* out) {
    constexpr bool MIXED = std::is_same<T, float>::value;
    const int prec = cg_prec(M);
    const size_t need = cg_work_layout(A, MIXED, prec, nullptr, nullptr);
    if (need > kNoAllocScratchBytes)
        throw Error(HSPS_ERR_NOMEM, "hsps_cg_solve_noalloc: the solve needs " + std::to_string(need) +
                                        " bytes of workspace, more than the " +
                                        std::to_string(kNoAllocScratchBytes) + "-byte scratch array");
    CgWork w;
    cg_work_layout(A, MIXED, prec, scratch_on(A.ctx->device), &w);
    CgState host[2];  // pageable: the state is read back synchronously (no events)
    w.host = host;
    return run_cg<T>(A, M, b, x, p, out, w);
}

// pipe: the pipelined variant (pipelined_fgmres_solver.cpp:56-243): CGS2
// orthogonalisation, breakdown ends the cycle, (its + 1) SpMVs charged per
// cycle.  Otherwise FGMRES with modified Gram-Schmidt (fgmres_solver.cpp).
template <class T>
static int run_fgmres(const Csr& A, const Precond* M, const double* b, double* x, const hsps_params& p,
                      hsps_stats* out, bool pipe = false) {
    constexpr bool MIXED = std::is_same<T, float>::value;
    Ctx* ctx = A.ctx;
    hipStream_t s = ctx->stream;
    const int n = A.rows;
    const double t0 = now_s();
    hsps_stats st{};
    if (MIXED && !A.v32.p && A.nnz) throw Error(HSPS_ERR_UNSUPPORTED, "matrix has no FP32 copy");
    const int m = std::max(1, std::min(p.restart_size, n));
    const bool has_m = M && M->kind != HSPS_PRECOND_NONE;
    const size_t ld = ((size_t)std::max(n, 1) + 63) / 64 * 64;
    DevBuf V(ctx, ld * (m + 1) * sizeof(T));
    DevBuf Z(ctx, has_m ? ld * m * sizeof(T) : 0);
    DevBuf r64(ctx, ld * sizeof(double));
    const int g = reduction_blocks(n);
    DevBuf partials(ctx, sizeof(double) * 2 * (size_t)std::max(std::max(g, A.n_blocks), 1));
    DevBuf dst(ctx, sizeof(GmState));
    DevBuf Hb(ctx, sizeof(double) * (size_t)(m + 1) * m);
    DevBuf csb(ctx, sizeof(double) * m), snb(ctx, sizeof(double) * m), gb(ctx, sizeof(double) * (m + 1)),
        yb(ctx, sizeof(double) * m);
    DevBuf scal(ctx, 64);
    DevBuf pipe_part(ctx, pipe ? sizeof(double) * (size_t)(m + 1) * (size_t)std::max(g, 1) : 0);
    DevBuf h2b(ctx, pipe ? sizeof(double) * (size_t)(m + 1) : 0);
    Pinned pin(sizeof(GmState) * 4 + 64);
    GmState* hst = static_cast<GmState*>(pin.p);
    GmState* d = dst.as<GmState>();
    double* part = partials.as<double>();
    double* H = Hb.as<double>();
    HSPS_HIP(hipMemsetAsync(Hb.p, 0, Hb.bytes, s));
    T* Vp = V.as<T>();
    T* Zp = has_m ? Z.as<T>() : Vp;
    auto Vk = [&](int k) { return Vp + (size_t)k * ld; };
    auto Zk = [&](int k) { return Zp + (size_t)k * ld; };

    // ||b|| (fgmres_solver.cpp:182-188)
    dot_f64_dev(n, b, b, scal.as<double>(), part, s);
    double bb = 0.0;
    HSPS_HIP(hipMemcpyAsync(&bb, scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HSPS_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(bb);
    auto finish = [&](int conv) {
        st.converged = conv;
        st.solve_time_s = now_s() - t0;
        if (out) *out = st;
        return conv;
    };
    if (bnorm < kEps) {
        st.final_residual = 0.0;
        return finish(1);
    }
    const double tol_abs = p.tol * bnorm;
    GmState init{};
    init.tol_abs = tol_abs;
    init.m = m;
    init.inner_rel = MIXED ? std::max(0.0, p.fgmres_inner_rel) : 0.0;
    HSPS_HIP(hipMemcpyAsync(d, &init, sizeof init, hipMemcpyHostToDevice, s));
    HSPS_HIP(hipStreamSynchronize(s));

    int total = 0, cycles = 0, conv = 0;
    double rel = 1.0;
    bool first = true;
    // The host reads the cycle state with one poll of lag, so up to two polls
    // of iterations are enqueued past the end of a cycle (their kernels exit
    // at once).  Expensive iterations (an AMG V-cycle, large operators) poll
    // (and a caller-written apply, whose calls past the end are wasted work)
    // every iteration; cheap ones every 4 to keep the launch queue ahead.
    const bool costly = (has_m && (M->kind == HSPS_PRECOND_AMG || M->kind == HSPS_PRECOND_USER)) || A.nnz > (1 << 21);
    const int steps_per_poll = costly ? 1 : 4;
    Event ev[2];
    while (total < p.max_iter) {
        const int cm = std::min(m, p.max_iter - total);
        residual_f64(A, b, x, r64.as<double>(), s, &d->beta2, part, nullptr);
        hipLaunchKernelGGL(k_gm_start<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r64.as<double>(), Vk(0), d, gb.as<double>());
        HSPS_CHECK_LAUNCH();
        int k = 0;
        bool pending = false;
        for (int j = 0; j < cm; ++j) {
            const int* skip = &d->cycle_done;
            if (has_m) {
                if (MIXED) precond_apply_f32(*M, (const float*)Vk(j), (float*)Zk(j), s, skip);
                else precond_apply_f64(*M, (const double*)Vk(j), (double*)Zk(j), s, skip);
            }
            T* w = Vk(j + 1);
            if (MIXED) spmv_f32(A, (const float*)Zk(j), (float*)w, s, skip);
            else spmv_f64(A, (const double*)Zk(j), (double*)w, s, skip);
            const OpGivens giv{j, H, csb.as<double>(), snb.as<double>(), gb.as<double>(), d};
            if (pipe) {
                const int nv = j + 1;
                double* pp = pipe_part.as<double>();
                double* h2 = h2b.as<double>();
                hipLaunchKernelGGL(k_cgs_dots<T>, dim3(g, nv), dim3(kBlock), 0, s, n, (const T*)w, (const T*)Vp, ld, skip,
                                   pp);
                hipLaunchKernelGGL(k_cgs_finish, dim3(1), dim3(kBlock), 0, s, nv, g, (const double*)pp, H + j, m, skip);
                hipLaunchKernelGGL((k_cgs_update<T, false>), dim3(g), dim3(kBlock), 0, s, n, w, (const T*)Vp, ld, nv,
                                   (const double*)(H + j), m, skip, pp);
                hipLaunchKernelGGL(k_cgs_dots<T>, dim3(g, nv), dim3(kBlock), 0, s, n, (const T*)w, (const T*)Vp, ld, skip,
                                   pp);
                hipLaunchKernelGGL(k_cgs_finish, dim3(1), dim3(kBlock), 0, s, nv, g, (const double*)pp, h2, 1, skip);
                hipLaunchKernelGGL((k_cgs_update<T, true>), dim3(g), dim3(kBlock), 0, s, n, w, (const T*)Vp, ld, nv,
                                   (const double*)h2, 1, skip, part);
                launch_finish<1>(part, g, OpGivensPipe{giv, h2}, s);
                HSPS_CHECK_LAUNCH();
            } else {
                for (int i = 0; i <= j; ++i) {
                    const double* hprev = i > 0 ? H + (size_t)(i - 1) * m + j : nullptr;
                    const T* vprev = i > 0 ? Vk(i - 1) : nullptr;
                    hipLaunchKernelGGL(k_mgs<T>, dim3(g), dim3(kBlock), 0, s, n, w, vprev, hprev, (const T*)Vk(i), skip,
                                       part);
                    launch_finish<1>(part, g, OpStoreH{H + (size_t)i * m + j, skip}, s);
                    HSPS_CHECK_LAUNCH();
                }
                hipLaunchKernelGGL(k_mgs_last<T>, dim3(g), dim3(kBlock), 0, s, n, (const double*)(H + (size_t)j * m + j),
                                   w, (const T*)Vk(j), (const GmState*)d, part);
                launch_finish<1>(part, g, giv, s);
            }
            hipLaunchKernelGGL(k_gm_scale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, w, (const GmState*)d);
            HSPS_CHECK_LAUNCH();
            if ((j + 1) % steps_per_poll == 0 && j + 1 < cm) {
                GmState* slot = hst + (k & 1);
                HSPS_HIP(hipMemcpyAsync(slot, d, sizeof(GmState), hipMemcpyDeviceToHost, s));
                HSPS_HIP(hipEventRecord(ev[k & 1].e, s));
                bool done = false;
                if (pending) {
                    HSPS_HIP(hipEventSynchronize(ev[(k - 1) & 1].e));
                    done = hst[(k - 1) & 1].cycle_done != 0;
                }
                pending = true;
                ++k;
                if (done) break;
            }
        }
        hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, (const double*)H, (const double*)gb.as<double>(),
                           yb.as<double>(), (const GmState*)d);
        hipLaunchKernelGGL(k_gm_xupdate<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, (const T*)Zp, ld,
                           (const double*)yb.as<double>(), (const GmState*)d);
        HSPS_CHECK_LAUNCH();
        HSPS_HIP(hipMemcpyAsync(hst + 2, d, sizeof(GmState), hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
        const GmState cur = hst[2];
        const int its = cur.jdone;
        if (first) { st.initial_residual = cur.beta; first = false; }
        total += its;
        ++cycles;
        rel = (its == 0 ? cur.beta : cur.res) / bnorm;
        if (MIXED) {
            // a cycle that starts below tol is the refinement loop's exit
            if (its == 0) { conv = rel < p.tol; break; }
        } else if (rel < p.tol || its == 0) {  // fgmres_solver.cpp:221-225
            conv = rel < p.tol;
            break;
        }
    }
    st.iterations = total;
    st.restarts = cycles - 1;
    st.final_residual = rel;
    st.true_residual = true_rel(A, b, x, r64.as<double>(), bnorm, scal.as<double>(), part, s);
    if (MIXED) conv = conv && st.true_residual < p.tol;
    // fgmres_solver.cpp:196-198: per inner iteration one SpMV and 4 n (vector accesses)
    const long long vb = MIXED ? 4 : 8;
    const long long spmvs = (long long)total + (pipe ? cycles : 0);  // pipelined_fgmres_solver.cpp:224-227
    finish_stats(&st, spmvs * 2LL * A.nnz + (long long)total * 4LL * n,
                 spmvs * model_spmv_bytes(n, A.nnz, vb) + (long long)total * 4LL * n * vb);
    return finish(conv);
}

}  // namespace hsps_dev

using namespace hsps_dev;

static void check_solve_args(hsps_csr A, hsps_precond M, const double* b, double* x) {
    if (!A || !b || !x) throw Error(HSPS_ERR_INVALID, "solve: null argument");
    if (A->rows != A->cols) throw Error(HSPS_ERR_SIZE, "solve: matrix must be square");
    if (M && M->n != A->rows) throw Error(HSPS_ERR_SIZE, "solve: preconditioner size != matrix size");
}

extern "C" {

int hsps_cg_solve(hsps_csr A, hsps_precond M, const double* b, double* x, const hsps_params* pp, hsps_stats* st) {
    return guarded([&] {
        check_solve_args(A, M, b, x);
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        A->ctx->set_current();
        // The engine rebuilds D^-1 from A, so it runs only for a Jacobi set
        // up from this very matrix and its current values; any other Jacobi
        // applies its own stored D^-1 through the plain kernels, as the
        // reference's CG applies M (cg_solver.cpp:79, 147).
        if (p.cg_engine && M && M->kind == HSPS_PRECOND_JACOBI && M->jacobi_omega == 1.0 && M->A0 == A &&
            M->A0_version == A->values_version) {
            const double t0 = now_s();
            SoloResult r;
            if (solo_cg(*A, b, x, p, &r)) {
                hsps_stats s{};
                s.iterations = r.iterations;
                s.converged = r.converged;
                s.restarts = r.restarts;
                s.reliable_updates = r.reliable_updates;
                s.initial_residual = r.initial_residual;
                s.final_residual = r.final_residual;
                s.true_residual = r.true_residual;
                s.solve_time_s = now_s() - t0;
                const CgCost cc = cg_cost(A->rows, A->nnz, r.iterations, r.converged != 0, r.restarts_pq,
                                          r.restarts - r.restarts_pq, p.precision == HSPS_MIXED ? 4 : 8);
                finish_stats(&s, cc.flops, cc.bytes);
                if (st) *st = s;
                return r.converged;
            }
        }
        const int rc = p.precision == HSPS_MIXED ? run_cg_alloc<float>(*A, M, b, x, p, st)
                                                 : run_cg_alloc<double>(*A, M, b, x, p, st);
        precond_check(M);
        return rc;
    });
}

int hsps_cg_solve_noalloc(hsps_csr A, hsps_precond M, const double* b, double* x, const hsps_params* pp,
                          hsps_stats* st) {
    return guarded([&] {
        check_solve_args(A, M, b, x);
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        A->ctx->set_current();
        const int rc = p.precision == HSPS_MIXED ? run_cg_noalloc<float>(*A, M, b, x, p, st)
                                                 : run_cg_noalloc<double>(*A, M, b, x, p, st);
        precond_check(M);
        return rc;
    });
}

int hsps_pipelined_fgmres_solve(hsps_csr A, hsps_precond M, const double* b, double* x, const hsps_params* pp,
                                hsps_stats* st) {
    return guarded([&] {
        check_solve_args(A, M, b, x);
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        A->ctx->set_current();
        const int rc = p.precision == HSPS_MIXED ? run_fgmres<float>(*A, M, b, x, p, st, true)
                                                 : run_fgmres<double>(*A, M, b, x, p, st, true);
        precond_check(M);
        return rc;
    });
}

int hsps_sstep_cg_solve(hsps_csr A, const double* b, double* x, const hsps_params* pp, hsps_stats* st) {
    return guarded([&] {
        check_solve_args(A, nullptr, b, x);
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        if (p.precision != HSPS_FP64) throw Error(HSPS_ERR_UNSUPPORTED, "s-step CG runs in FP64 only");
        A->ctx->set_current();
        return sstep_cg_dev(*A, b, x, p, st);
    });
}

int hsps_spectral_interval(hsps_csr A, int steps, double* lo, double* hi) {
    return guarded([&] {
        if (!A || !lo || !hi) throw Error(HSPS_ERR_INVALID, "hsps_spectral_interval: null argument");
        if (A->rows != A->cols) throw Error(HSPS_ERR_SIZE, "spectral_interval: matrix must be square");
        A->ctx->set_current();
        spectral_interval_dev(*A, steps, lo, hi);
        return HSPS_OK;
    });
}

int hsps_fgmres_solve(hsps_csr A, hsps_precond M, const double* b, double* x, const hsps_params* pp,
                      hsps_stats* st) {
    return guarded([&] {
        check_solve_args(A, M, b, x);
        hsps_params p;
        if (pp) p = *pp; else hsps_params_default(&p);
        A->ctx->set_current();
        const int rc = p.precision == HSPS_MIXED ? run_fgmres<float>(*A, M, b, x, p, st)
                                                 : run_fgmres<double>(*A, M, b, x, p, st);
        precond_check(M);
        return rc;
    });
}

}  // extern "C"
