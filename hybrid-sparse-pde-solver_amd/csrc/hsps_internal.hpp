// =============================================================================
// hsps_internal.hpp — host-side internals of the gfx950 solve path: handle
// structs, error plumbing, workspace helpers.  Not part of the C ABI.
// =============================================================================
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hsps_dev.h"

namespace hsps_dev {

// ---- error plumbing: exceptions inside, status codes at the boundary --------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// A failed call also clears HIP's last-error slot, so that a later launch
// check does not report this (already thrown) error again.
#define HSPS_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            (void)hipGetLastError();                                                              \
            throw ::hsps_dev::Error(HSPS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                                         \
    } while (0)

#define HSPS_CHECK_LAUNCH() HSPS_HIP(hipGetLastError())

// Runs body; maps exceptions to status codes and the thread-local message.
template <class F>
int guarded(F&& body) {
    try {
        return body();
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HSPS_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return HSPS_ERR_INVALID;
    }
}

// ---- device buffer with RAII, tracked by the owning context -------------------
struct Ctx;
struct DevBuf {
    Ctx* ctx = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(Ctx* c, size_t b);
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept;
    void reset();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// A context lives until hsps_ctx_destroy has been called AND every handle or
// hsps_malloc allocation made through it has been released, so handles may
// be destroyed in any order (e.g. by a garbage collector at exit).
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    std::atomic<size_t> bytes_in_use{0};
    std::atomic<int> refs{1};
    // scratch for standalone BLAS-1 reductions (stream ordered)
    DevBuf red_partials, red_out;
    double* pinned = nullptr;  // host pinned readback area (64 doubles)
    void set_current() const { HSPS_HIP(hipSetDevice(device)); }
};
void ctx_retain(Ctx* c);
void ctx_release(Ctx* c);

// Host -> device copies and fills ordered on a stream of ours, returning once
// done.  Never through the legacy null stream: it orders against every
// blocking stream of the process, and HIP refuses it while any other thread
// captures a graph.  Pageable sources must stay alive until the call returns.
inline void upload_on(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    HSPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HSPS_HIP(hipStreamSynchronize(s));
}
inline void fill_on(hipStream_t s, void* dst, int value, size_t bytes) {
    if (!bytes) return;
    HSPS_HIP(hipMemsetAsync(dst, value, bytes, s));
    HSPS_HIP(hipStreamSynchronize(s));
}

// SpMV schedule constants (gfx950: 64-wide waves, 160 KB LDS per CU).
constexpr int kBlock = 256;       // threads per workgroup (4 waves)
constexpr int kLdsNnz = 2048;     // FP64 products staged in LDS per row block (16 KB)
constexpr int kMaxBlockRows = 256;
constexpr int kLdsNnz32 = 4096;   // FP32 products in the same 16 KB
constexpr int kMaxBlockRows32 = 512;
constexpr int kMaxRedBlocks = 2048;  // cap for grid-stride reduction kernels

// Device CSR: int32 pattern, FP64 values + FP32 copy, diagonal, row blocks.
struct Batch;  // batch.hip

struct Csr {
    Ctx* ctx = nullptr;
    int rows = 0, cols = 0, nnz = 0;
    std::vector<int> h_rp, h_ci;
    std::vector<double> h_v;
    DevBuf rp, ci, v64, v32, diag64, blocks, blocks32;
    int n_blocks = 0;    // FP64 row blocks (kMaxBlockRows / kLdsNnz)
    int n_blocks32 = 0;  // FP32 row blocks (kMaxBlockRows32 / kLdsNnz32), <= n_blocks
    // this matrix as a one-system batch (built on the first PCG + Jacobi
    // solve that uses the batched engine; dropped when the values change)
    std::shared_ptr<Batch> solo;
    int solo_use = -1;  // -1 not decided, 0 plain kernels, 1 batched engine
    // bumped by every hsps_csr_update_values: a preconditioner records the
    // version it was set up from
    unsigned long long values_version = 0;
    void upload_values();
    void build_schedule();
};

// PCG + Jacobi (omega = 1) on one system through the batched engine
// (batch.hip): diagonal / stencil storage and the fused iteration kernels.
struct SoloResult {
    int iterations = 0, converged = 0, restarts = 0, reliable_updates = 0;
    int restarts_pq = 0;  // restarts after a p.q breakdown (the others follow a rho breakdown)
    double initial_residual = 0, final_residual = 0, true_residual = 0;
};
// Returns false (nothing run) for a large CSR-stored system: the plain
// kernel sequence serves those better.
bool solo_cg(Csr& A, const double* b, double* x, const hsps_params& p, SoloResult* out);

// flop_model (energy_monitor.hpp:163-177) with vb-byte values and vectors
// (8: the reference's Real; 4: the FP32 inner loop of mixed precision).
inline long long model_spmv_bytes(long long rows, long long nnz, long long vb) {
    return (rows + 1 + nnz) * 4 + nnz * vb + rows * vb * 2;
}
// CGSolver::solve's counters (cg_solver.cpp:98-175): a full iteration is an
// SpMV, p.q, two axpys, r.z and the p update (2 nnz + 10 n FLOPs, SpMV bytes
// + 13 n vector accesses); the converging iteration stops after the axpys
// (2 nnz + 6 n, + 8 n); a p.q breakdown after p.q (2 nnz + 2 n, + 2 n); a
// rho breakdown after r.z (2 nnz + 8 n, + 10 n).  Breakdown iterations count
// as iterations (the loop's `continue`).  The residual norm is not counted.
struct CgCost {
    long long flops = 0, bytes = 0;
};
inline CgCost cg_cost(long long n, long long nnz, long long iters, bool converged_last, long long pq_breaks,
                      long long rho_breaks, long long vb) {
    converged_last = converged_last && iters > 0;  // 0 iterations: converged at the start
    const long long full = iters - (converged_last ? 1 : 0) - pq_breaks - rho_breaks;
    const long long S = model_spmv_bytes(n, nnz, vb);
    CgCost c;
    c.flops = full * (2 * nnz + 10 * n) + (converged_last ? 2 * nnz + 6 * n : 0) + pq_breaks * (2 * nnz + 2 * n) +
              rho_breaks * (2 * nnz + 8 * n);
    c.bytes = full * (S + 13 * n * vb) + (converged_last ? S + 8 * n * vb : 0) + pq_breaks * (S + 2 * n * vb) +
              rho_breaks * (S + 10 * n * vb);
    return c;
}

// Row-block partition: consecutive rows, <= max_rows rows and <= max_nnz
// nonzeros per block; a longer row gets a block of its own.  Greedy, so a
// looser (max_rows, max_nnz) never yields more blocks.
std::vector<int> make_row_blocks(const int* rp, int rows, int max_rows, int max_nnz);

// Wrap a device CSR from host arrays (used by ILU factors).
std::unique_ptr<Csr> make_csr(Ctx* ctx, int rows, int cols, std::vector<int> rp, std::vector<int> ci,
                              std::vector<double> v, bool want_f32 = true);
// Take over a CSR built on the device (AMG levels): h_rp is the row pointer
// already copied to the host; h_ci / h_v stay empty.  spmv_ready adds the
// FP32 values, the diagonal and the row-block schedule.
std::unique_ptr<Csr> csr_adopt(Ctx* ctx, int rows, int cols, DevBuf&& rp, DevBuf&& ci, DevBuf&& v64,
                               std::vector<int>&& h_rp, bool spmv_ready);

// ---- launchers implemented in the .hip files ---------------------------------
// SpMV variants (spmv.hip)
void spmv_f64(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip = nullptr);
void spmv_f32(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip = nullptr);
// r = b - A x (FP64), and optionally ||r||^2 into *d_nrm2 (device scalar;
// needs a partials buffer of >= A.n_blocks doubles).
void residual_f64(const Csr& A, const double* b, const double* x, double* r, hipStream_t s,
                  double* d_nrm2 = nullptr, double* partials = nullptr, const int* skip = nullptr);
// y += A x (FP64 or FP32), y = A x as the 'add' flag says
void spmv_add_f64(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip);
void spmv_add_f32(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip);
// xo = xi + omega * dinv .* (rhs - A xi)
void smooth_f64(const Csr& A, const double* dinv, double omega, const double* rhs, const double* xi,
                double* xo, hipStream_t s, const int* skip);
void smooth_f32(const Csr& A, const float* dinv, float omega, const float* rhs, const float* xi, float* xo,
                hipStream_t s, const int* skip);
void residual_f32(const Csr& A, const float* b, const float* x, float* r, hipStream_t s, const int* skip);

// BLAS-1 (blas.hip)
// x.y into *d_out (device scalar); partials >= reduction_blocks(n) doubles.
void dot_f64_dev(int n, const double* x, const double* y, double* d_out, double* partials, hipStream_t s);
void axpy_f64(int n, double a, const double* x, double* y, hipStream_t s);
void axpby_f64(int n, double a, const double* x, double b, double* y, hipStream_t s);
void scale_f64(int n, double a, double* x, hipStream_t s);
void copy_f64_to_f32(int n, const double* x, float* y, hipStream_t s, const int* skip = nullptr);
void fill_zero_f64(int n, double* x, hipStream_t s, const int* skip = nullptr);
void fill_zero_f32(int n, float* x, hipStream_t s, const int* skip = nullptr);
int reduction_blocks(int n);

// s-step CG and its Lanczos spectral interval (sstep.hip)
void spectral_interval_dev(const Csr& A, int steps, double* lo, double* hi);
int sstep_cg_dev(const Csr& A, const double* b, double* x, const hsps_params& p, hsps_stats* out);

// Preconditioner object (precond.hip)
struct Precond;
void precond_apply_f64(const Precond& M, const double* r, double* z, hipStream_t s, const int* skip);
void precond_apply_f32(const Precond& M, const float* r, float* z, hipStream_t s, const int* skip);
// Raises the error a preconditioner's sweeps flagged on the device since the
// last check (synchronises the context's stream when there is one to read).
void precond_check(const Precond* M);
int precond_kind(const Precond& M);
const double* precond_jacobi_dinv64(const Precond& M);
const float* precond_jacobi_dinv32(const Precond& M);

}  // namespace hsps_dev

struct hsps_ctx_s : hsps_dev::Ctx {};
// A matrix handle lives until hsps_csr_destroy has been called AND every
// preconditioner built on it has been destroyed (they read it at apply time).
struct hsps_csr_s : hsps_dev::Csr {
    std::atomic<int> refs{1};
};
namespace hsps_dev {
void csr_retain(hsps_csr_s* A);
void csr_release(hsps_csr_s* A);
}  // namespace hsps_dev
