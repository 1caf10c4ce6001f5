// =============================================================================
// core.hip — context, device memory, CSR upload + row-block schedule, SpMV and
// BLAS-1 entry points of the C ABI.
//
// Reference counterparts: SparseMatrix (sparse_matrix.hpp:34-170,
// sparse_matrix.cpp:68-87), Vector BLAS-1 (vector.cpp:39-119), and the role of
// the CUDA backend's DeviceSparseMatrix / cusparseSpMV / cublasD*
// (cuda_backend.cpp:26-431) — here with everything resident on the device.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>

#include "device_common.hpp"

namespace hsps_dev {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

void ctx_retain(Ctx* c) { c->refs.fetch_add(1); }
void dev_release_all(Ctx* c);
void ctx_release(Ctx* c) {
    if (c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->red_partials.reset();
    c->red_out.reset();
    dev_release_all(c);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete static_cast<hsps_ctx_s*>(c);
}

void csr_retain(hsps_csr_s* A) { A->refs.fetch_add(1); }
void csr_release(hsps_csr_s* A) {
    if (A->refs.fetch_sub(1) != 1) return;
    Ctx* c = A->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete A;
    ctx_release(c);
}

// ---- device memory cache -------------------------------------------------------
// hipFree waits for the work of every stream on the device, so a solve that
// frees its buffers while other contexts' streams are busy (a request server
// has one context per worker) would wait for all of their kernels.  Freed
// blocks are cached instead:
//  * a block a context frees stays with that context, tagged with an event
//    recorded on its stream, and is handed out again (same size class) to
//    that context once the event has completed, i.e. once everything the
//    context enqueued before freeing it has run.  Only the owning context
//    queries its blocks' events, so no thread ever touches another thread's
//    stream (a query of an event whose stream is being captured into a graph
//    would disturb the capture);
//  * blocks known to be idle (hsps_free drains the stream first; a destroyed
//    context's blocks once its stream is drained) go to a per-device pool
//    any context may take from.
// Work on any other stream that used a block must be complete, or ordered
// before that point, when it is freed (the library's second streams are
// joined into the context's stream before their buffers are released;
// hsps_ctx_set_stream drains the previous stream).  Sizes are rounded up to
// classes 2^k and 3 * 2^(k-1) (at most a third wasted).  The cache keeps at
// most HSPS_MEM_CACHE_MB per device (default: a quarter of the device's
// memory; 0 turns it off) and gives blocks back with hipFree when an
// allocation fails.
namespace {
struct CachedBlock {
    void* p;
    hipEvent_t ev;
};
struct MemCache {
    std::mutex mu;
    std::map<std::pair<const Ctx*, size_t>, std::deque<CachedBlock>> owned;  // (context, class), oldest first
    std::map<std::pair<int, size_t>, std::vector<void*>> idle;               // (device, class)
    std::map<int, std::vector<hipEvent_t>> events;                           // completed events, per device
    std::unordered_map<void*, size_t> user;                                  // hsps_malloc'd pointer -> bytes
    std::map<int, size_t> bytes, cap;                                        // cached bytes and limit, per device
};
MemCache& mem_cache() {
    static MemCache* m = new MemCache;  // never destroyed: cached blocks live until the process ends
    return *m;
}
size_t size_class(size_t b) {
    size_t c = 256;
    while (c < b) {
        const size_t h = c + c / 2;
        if (h >= b) return h;
        c *= 2;
    }
    return c;
}
size_t cache_cap(MemCache& M, int dev) {  // M.mu held, dev current
    auto it = M.cap.find(dev);
    if (it != M.cap.end()) return it->second;
    size_t cap = 0;
    if (const char* e = std::getenv("HSPS_MEM_CACHE_MB")) {
        cap = (size_t)std::max(0LL, std::atoll(e)) << 20;
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) cap = tot / 4;
        else (void)hipGetLastError();
    }
    M.cap[dev] = cap;
    return cap;
}
// Gives idle blocks of the device back (hipFree) until at most keep bytes
// stay cached; with c, also c's own blocks whose event has completed (all of
// them, waiting for their events, when all is set).
void cache_trim(MemCache& M, int dev, const Ctx* c, size_t keep, bool all) {  // M.mu held, dev current
    for (auto it = M.idle.begin(); it != M.idle.end() && M.bytes[dev] > keep;) {
        if (it->first.first != dev) { ++it; continue; }
        auto& v = it->second;
        while (!v.empty() && M.bytes[dev] > keep) {
            (void)hipFree(v.back());
            v.pop_back();
            M.bytes[dev] -= it->first.second;
        }
        it = v.empty() ? M.idle.erase(it) : std::next(it);
    }
    if (!c) return;
    for (auto it = M.owned.begin(); it != M.owned.end() && M.bytes[dev] > keep;) {
        if (it->first.first != c) { ++it; continue; }
        auto& q = it->second;
        while (!q.empty() && M.bytes[dev] > keep) {
            CachedBlock b = q.front();
            if (all) {
                (void)hipEventSynchronize(b.ev);
            } else if (hipEventQuery(b.ev) != hipSuccess) {
                (void)hipGetLastError();  // not ready: keep it
                break;
            }
            q.pop_front();
            M.events[dev].push_back(b.ev);
            (void)hipFree(b.p);
            M.bytes[dev] -= it->first.second;
        }
        it = q.empty() ? M.owned.erase(it) : std::next(it);
    }
}
}  // namespace

// A block of at least b bytes for c: one of c's cached blocks whose last
// user has finished, else an idle one, else a new one.
void* dev_alloc(Ctx* c, size_t b) {
    c->set_current();
    MemCache& M = mem_cache();
    const size_t cls = size_class(b);
    {
        std::lock_guard<std::mutex> lk(M.mu);
        auto it = M.owned.find({c, cls});
        if (it != M.owned.end()) {
            auto& q = it->second;
            for (size_t k = 0; k < q.size() && k < 8; ++k) {  // oldest first: the likeliest done
                if (hipEventQuery(q[k].ev) != hipSuccess) {
                    (void)hipGetLastError();  // not ready
                    continue;
                }
                void* p = q[k].p;
                M.events[c->device].push_back(q[k].ev);
                q.erase(q.begin() + (long)k);
                M.bytes[c->device] -= cls;
                return p;
            }
        }
        auto jt = M.idle.find({c->device, cls});
        if (jt != M.idle.end() && !jt->second.empty()) {
            void* p = jt->second.back();
            jt->second.pop_back();
            M.bytes[c->device] -= cls;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(M.mu);
            cache_trim(M, c->device, c, 0, true);
        }
        e = hipMalloc(&p, cls);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(HSPS_ERR_NOMEM, "hipMalloc(" + std::to_string(b) + " B) failed: " + hipGetErrorString(e));
        }
    }
    return p;
}

// Returns a block of b bytes from dev_alloc to the cache: c's own, behind the
// work already enqueued on c's stream, or the idle pool when that work is
// known to be done.
void dev_release(Ctx* c, void* p, size_t b, bool idle) {
    if (!p) return;
    c->set_current();
    MemCache& M = mem_cache();
    const size_t cls = size_class(b);
    std::lock_guard<std::mutex> lk(M.mu);
    const size_t cap = cache_cap(M, c->device);
    if (cap < cls) {  // no caching: the plain free
        (void)hipFree(p);
        return;
    }
    if (idle) {
        M.idle[{c->device, cls}].push_back(p);
    } else {
        hipEvent_t ev = nullptr;
        auto& evs = M.events[c->device];
        if (!evs.empty()) {
            ev = evs.back();
            evs.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ev = nullptr;
        }
        if (!ev || hipEventRecord(ev, c->stream) != hipSuccess) {
            (void)hipGetLastError();
            if (ev) evs.push_back(ev);
            (void)hipFree(p);
            return;
        }
        M.owned[{c, cls}].push_back({p, ev});
    }
    M.bytes[c->device] += cls;
    if (M.bytes[c->device] > cap) cache_trim(M, c->device, c, cap, false);
}

// A context going away: its stream is drained, so its cached blocks become
// idle for every other context.
void dev_release_all(Ctx* c) {
    MemCache& M = mem_cache();
    std::lock_guard<std::mutex> lk(M.mu);
    for (auto it = M.owned.begin(); it != M.owned.end();) {
        if (it->first.first != c) { ++it; continue; }
        for (const CachedBlock& b : it->second) {
            (void)hipEventSynchronize(b.ev);
            M.events[c->device].push_back(b.ev);
            M.idle[{c->device, it->first.second}].push_back(b.p);
        }
        it = M.owned.erase(it);
    }
}

// ---- DevBuf ------------------------------------------------------------------
DevBuf::DevBuf(Ctx* c, size_t b) : ctx(c), bytes(b) {
    if (b == 0) return;
    p = dev_alloc(c, b);
    c->bytes_in_use += b;
}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
    if (this != &o) {
        reset();
        ctx = o.ctx; p = o.p; bytes = o.bytes;
        o.p = nullptr; o.bytes = 0;
    }
    return *this;
}
void DevBuf::reset() {
    if (p) {
        if (ctx) {
            dev_release(ctx, p, bytes, false);
            ctx->bytes_in_use -= bytes;
        } else {
            (void)hipFree(p);
        }
    }
    p = nullptr;
    bytes = 0;
}
DevBuf::~DevBuf() { reset(); }

// ---- row-block schedule ------------------------------------------------------
std::vector<int> make_row_blocks(const int* rp, int rows, int max_rows, int max_nnz) {
    std::vector<int> b;
    b.reserve(rows / 64 + 2);
    b.push_back(0);
    int start = 0;
    while (start < rows) {
        int end = start;
        const int base = rp[start];
        if (rp[start + 1] - base > max_nnz) {
            end = start + 1;  // long row: own block
        } else {
            while (end < rows && end - start < max_rows && rp[end + 1] - base <= max_nnz) ++end;
        }
        b.push_back(end);
        start = end;
    }
    return b;
}

void Csr::build_schedule() {
    std::vector<int> b = make_row_blocks(h_rp.data(), rows, kMaxBlockRows, kLdsNnz);
    n_blocks = (int)b.size() - 1;
    blocks = DevBuf(ctx, b.size() * sizeof(int));
    upload_on(ctx->stream, blocks.p, b.data(), b.size() * sizeof(int));
    std::vector<int> b32 = make_row_blocks(h_rp.data(), rows, kMaxBlockRows32, kLdsNnz32);
    n_blocks32 = (int)b32.size() - 1;
    // partial-sum buffers are sized by n_blocks for both precisions
    if (n_blocks32 > n_blocks) throw Error(HSPS_ERR_HIP, "csr: FP32 schedule has more blocks than FP64");
    blocks32 = DevBuf(ctx, b32.size() * sizeof(int));
    upload_on(ctx->stream, blocks32.p, b32.data(), b32.size() * sizeof(int));
}

__global__ void k_f64_to_f32(int n, const double* __restrict__ x, float* __restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = (float)x[i];
}

void Csr::upload_values() {
    ctx->set_current();
    solo.reset();  // a one-system batch holds its own copy of the values
    solo_use = -1;
    hipStream_t s = ctx->stream;  // stream-ordered: no device-wide sync (other contexts keep running)
    std::vector<double> d(rows, 0.0);
    for (int i = 0; i < rows; ++i)
        for (int k = h_rp[i]; k < h_rp[i + 1]; ++k)
            if (h_ci[k] == i) { d[i] = h_v[k]; break; }
    if (nnz > 0) {
        HSPS_HIP(hipMemcpyAsync(v64.p, h_v.data(), (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, s));
        if (v32.p) {
            hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, v64.as<double>(),
                               v32.as<float>());
            HSPS_CHECK_LAUNCH();
        }
    }
    if (rows > 0) HSPS_HIP(hipMemcpyAsync(diag64.p, d.data(), (size_t)rows * sizeof(double), hipMemcpyHostToDevice, s));
    HSPS_HIP(hipStreamSynchronize(s));
}

static void validate_csr(int rows, int cols, int nnz, const int* rp, const int* ci) {
    if (rows < 0 || cols < 0 || nnz < 0) throw Error(HSPS_ERR_INVALID, "csr: negative dimension");
    if (rp[0] != 0 || rp[rows] != nnz) throw Error(HSPS_ERR_INVALID, "csr: row_ptr[0] must be 0 and row_ptr[rows] == nnz");
    for (int i = 0; i < rows; ++i)
        if (rp[i + 1] < rp[i]) throw Error(HSPS_ERR_INVALID, "csr: row_ptr not monotone at row " + std::to_string(i));
    for (int k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= cols) throw Error(HSPS_ERR_INVALID, "csr: column index out of range at entry " + std::to_string(k));
}

// Raw diagonal: the first entry with column == row, 0 when absent.
__global__ void k_raw_diag(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                           double* __restrict__ d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                a = v[k];
                break;
            }
        d[i] = a;
    }
}

std::unique_ptr<Csr> csr_adopt(Ctx* ctx, int rows, int cols, DevBuf&& rp, DevBuf&& ci, DevBuf&& v64,
                               std::vector<int>&& h_rp, bool spmv_ready) {
    auto A = std::make_unique<Csr>();
    A->ctx = ctx;
    A->rows = rows;
    A->cols = cols;
    A->nnz = h_rp.empty() ? 0 : h_rp[rows];
    A->h_rp = std::move(h_rp);
    A->rp = std::move(rp);
    A->ci = std::move(ci);
    A->v64 = std::move(v64);
    if (spmv_ready) {
        ctx->set_current();
        hipStream_t s = ctx->stream;
        A->v32 = DevBuf(ctx, (size_t)A->nnz * sizeof(float));
        A->diag64 = DevBuf(ctx, (size_t)rows * sizeof(double));
        if (A->nnz) {
            hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(A->nnz)), dim3(kBlock), 0, s, A->nnz, A->v64.as<double>(),
                               A->v32.as<float>());
            HSPS_CHECK_LAUNCH();
        }
        if (rows) {
            hipLaunchKernelGGL(k_raw_diag, dim3(grid_for(rows, 1)), dim3(kBlock), 0, s, rows, A->rp.as<int>(),
                               A->ci.as<int>(), A->v64.as<double>(), A->diag64.as<double>());
            HSPS_CHECK_LAUNCH();
        }
        HSPS_HIP(hipStreamSynchronize(s));
        A->build_schedule();
    }
    return A;
}

std::unique_ptr<Csr> make_csr(Ctx* ctx, int rows, int cols, std::vector<int> rp, std::vector<int> ci,
                              std::vector<double> v, bool want_f32) {
    validate_csr(rows, cols, (int)v.size(), rp.data(), ci.data());
    auto A = std::make_unique<Csr>();
    A->ctx = ctx;
    A->rows = rows;
    A->cols = cols;
    A->nnz = (int)v.size();
    A->h_rp = std::move(rp);
    A->h_ci = std::move(ci);
    A->h_v = std::move(v);
    A->rp = DevBuf(ctx, (size_t)(rows + 1) * sizeof(int));
    A->ci = DevBuf(ctx, (size_t)A->nnz * sizeof(int));
    A->v64 = DevBuf(ctx, (size_t)A->nnz * sizeof(double));
    if (want_f32) A->v32 = DevBuf(ctx, (size_t)A->nnz * sizeof(float));
    A->diag64 = DevBuf(ctx, (size_t)rows * sizeof(double));
    ctx->set_current();
    upload_on(ctx->stream, A->rp.p, A->h_rp.data(), (size_t)(rows + 1) * sizeof(int));
    upload_on(ctx->stream, A->ci.p, A->h_ci.data(), (size_t)A->nnz * sizeof(int));
    A->upload_values();
    A->build_schedule();
    return A;
}

// ---- SpMV launchers -------------------------------------------------------------
void spmv_f64(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip) {
    EpiStore<double> e;
    e.flag = skip;
    e.y = y;
    launch_spmv(A, A.v64.as<double>(), x, e, s);
}
void spmv_f32(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip) {
    EpiStore<float> e;
    e.flag = skip;
    e.y = y;
    launch_spmv(A, A.v32.as<float>(), x, e, s);
}
void spmv_add_f64(const Csr& A, const double* x, double* y, hipStream_t s, const int* skip) {
    EpiAdd<double> e;
    e.flag = skip;
    e.y = y;
    launch_spmv(A, A.v64.as<double>(), x, e, s);
}
void spmv_add_f32(const Csr& A, const float* x, float* y, hipStream_t s, const int* skip) {
    EpiAdd<float> e;
    e.flag = skip;
    e.y = y;
    launch_spmv(A, A.v32.as<float>(), x, e, s);
}
void residual_f64(const Csr& A, const double* b, const double* x, double* r, hipStream_t s, double* d_nrm2,
                  double* partials, const int* skip) {
    if (d_nrm2) {
        EpiResidual<double, true> e;
        e.flag = skip; e.b = b; e.r = r; e.partials = partials;
        launch_spmv(A, A.v64.as<double>(), x, e, s);
        OpStore<1> op;
        op.out = d_nrm2;
        op.flag = skip;
        launch_finish<1>(partials, A.n_blocks, op, s);
        HSPS_CHECK_LAUNCH();
    } else {
        EpiResidual<double, false> e;
        e.flag = skip; e.b = b; e.r = r;
        launch_spmv(A, A.v64.as<double>(), x, e, s);
    }
}
void residual_f32(const Csr& A, const float* b, const float* x, float* r, hipStream_t s, const int* skip) {
    EpiResidual<float, false> e;
    e.flag = skip; e.b = b; e.r = r;
    launch_spmv(A, A.v32.as<float>(), x, e, s);
}
void smooth_f64(const Csr& A, const double* dinv, double omega, const double* rhs, const double* xi, double* xo,
                hipStream_t s, const int* skip) {
    EpiSmooth<double> e;
    e.flag = skip; e.dinv = dinv; e.rhs = rhs; e.xi = xi; e.xo = xo; e.omega = omega;
    launch_spmv(A, A.v64.as<double>(), xi, e, s);
}
void smooth_f32(const Csr& A, const float* dinv, float omega, const float* rhs, const float* xi, float* xo,
                hipStream_t s, const int* skip) {
    EpiSmooth<float> e;
    e.flag = skip; e.dinv = dinv; e.rhs = rhs; e.xi = xi; e.xo = xo; e.omega = omega;
    launch_spmv(A, A.v32.as<float>(), xi, e, s);
}

// ---- BLAS-1 kernels ------------------------------------------------------------
int reduction_blocks(int n) { return std::min(grid_for(n, 8), kMaxRedBlocks); }

__global__ void __launch_bounds__(kBlock)
k_dot(int n, const double* __restrict__ x, const double* __restrict__ y, double* partials) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[0] += x[i] * y[i];
    block_partial<1>(v, partials);
}
void dot_f64_dev(int n, const double* x, const double* y, double* d_out, double* partials, hipStream_t s) {
    const int g = reduction_blocks(n);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(kBlock), 0, s, n, x, y, partials);
    OpStore<1> op;
    op.out = d_out;
    launch_finish<1>(partials, g, op, s);
    HSPS_CHECK_LAUNCH();
}

__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y, int mode) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (mode == 0) y[i] += a * x[i];          // axpy
        else if (mode == 1) y[i] = a * x[i] + b * y[i];  // axpby
        else y[i] *= a;                                    // scale
    }
}
void axpy_f64(int n, double a, const double* x, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, 0.0, y, 0);
    HSPS_CHECK_LAUNCH();
}
void axpby_f64(int n, double a, const double* x, double b, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, b, y, 1);
    HSPS_CHECK_LAUNCH();
}
void scale_f64(int n, double a, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, nullptr, 0.0, x, 2);
    HSPS_CHECK_LAUNCH();
}

__global__ void k_cvt(int n, const double* __restrict__ x, float* __restrict__ y, const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = (float)x[i];
}
void copy_f64_to_f32(int n, const double* x, float* y, hipStream_t s, const int* skip) {
    hipLaunchKernelGGL(k_cvt, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, y, skip);
    HSPS_CHECK_LAUNCH();
}
template <class T>
__global__ void k_zero(int n, T* __restrict__ x, const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = T(0);
}
void fill_zero_f64(int n, double* x, hipStream_t s, const int* skip) {
    hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, skip);
    HSPS_CHECK_LAUNCH();
}
void fill_zero_f32(int n, float* x, hipStream_t s, const int* skip) {
    hipLaunchKernelGGL(k_zero<float>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, skip);
    HSPS_CHECK_LAUNCH();
}

}  // namespace hsps_dev

using namespace hsps_dev;

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

const char* hsps_last_error(void) { return g_last_error.c_str(); }

void hsps_params_default(hsps_params* p) {
    if (!p) return;
    p->tol = 1e-8;
    p->max_iter = 2000;
    p->restart_size = 50;
    p->drop_tol = 1e-4;
    p->fill_per_row = 0;
    p->amg_levels = 10;
    p->amg_strong = 0.25;
    p->amg_smooth_pre = 2;
    p->amg_smooth_post = 2;
    p->amg_omega = 0.67;
    p->jacobi_omega = 1.0;
    p->precision = HSPS_FP64;
    p->reliable_delta = 0.1;
    p->check_every = 32;
    p->batch_tile_rows = 0;
    p->batch_graphs = 1;
    p->cg_engine = 1;
    p->fgmres_inner_rel = 0.0;
    p->sstep_k = 1;
    p->sstep_cond_tol = 1e12;
}

int hsps_ctx_create(int device, hsps_ctx* out) {
    return guarded([&] {
        if (!out) throw Error(HSPS_ERR_INVALID, "hsps_ctx_create: out is NULL");
        int n = 0;
        HSPS_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n)
            throw Error(HSPS_ERR_INVALID, "hsps_ctx_create: device " + std::to_string(device) + " not present (" +
                                              std::to_string(n) + " devices)");
        auto* c = new hsps_ctx_s;
        c->device = device;
        try {
            HSPS_HIP(hipSetDevice(device));
            HSPS_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
            c->stream = c->own_stream;
            c->red_partials = DevBuf(c, sizeof(double) * kMaxRedBlocks * 4);
            c->red_out = DevBuf(c, 256);
            HSPS_HIP(hipHostMalloc((void**)&c->pinned, 64 * sizeof(double), hipHostMallocDefault));
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
        return HSPS_OK;
    });
}

int hsps_ctx_destroy(hsps_ctx ctx) {
    return guarded([&] {
        if (ctx) ctx_release(ctx);
        return HSPS_OK;
    });
}

int hsps_ctx_set_stream(hsps_ctx ctx, void* stream) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        ctx->set_current();
        // blocks freed from now on are fenced on the new stream only: the work
        // enqueued on the previous one completes first
        HSPS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
        return HSPS_OK;
    });
}
void* hsps_ctx_stream(hsps_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }
int hsps_ctx_sync(hsps_ctx ctx) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        HSPS_HIP(hipStreamSynchronize(ctx->stream));
        return HSPS_OK;
    });
}
size_t hsps_ctx_bytes_in_use(hsps_ctx ctx) { return ctx ? ctx->bytes_in_use.load() : size_t(0); }

int hsps_malloc(hsps_ctx ctx, size_t bytes, void** d_ptr) {
    return guarded([&] {
        if (!ctx || !d_ptr) throw Error(HSPS_ERR_INVALID, "hsps_malloc: null argument");
        ctx->set_current();
        *d_ptr = nullptr;
        if (bytes == 0) return HSPS_OK;
        void* p = dev_alloc(ctx, bytes);
        {
            MemCache& M = mem_cache();
            std::lock_guard<std::mutex> lk(M.mu);
            M.user[p] = bytes;
        }
        *d_ptr = p;
        ctx_retain(ctx);
        return HSPS_OK;
    });
}
int hsps_free(hsps_ctx ctx, void* d_ptr) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        if (d_ptr) {
            ctx->set_current();
            HSPS_HIP(hipStreamSynchronize(ctx->stream));
            size_t bytes = 0;
            {
                MemCache& M = mem_cache();
                std::lock_guard<std::mutex> lk(M.mu);
                auto it = M.user.find(d_ptr);
                if (it != M.user.end()) {
                    bytes = it->second;
                    M.user.erase(it);
                }
            }
            if (bytes) dev_release(ctx, d_ptr, bytes, true);  // the stream is drained
            else HSPS_HIP(hipFree(d_ptr));
            ctx_release(ctx);
        }
        return HSPS_OK;
    });
}
int hsps_memcpy_h2d(hsps_ctx ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        if (bytes == 0) return HSPS_OK;
        ctx->set_current();
        HSPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        HSPS_HIP(hipStreamSynchronize(ctx->stream));
        return HSPS_OK;
    });
}
int hsps_memcpy_d2h(hsps_ctx ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        if (bytes == 0) return HSPS_OK;
        ctx->set_current();
        HSPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HSPS_HIP(hipStreamSynchronize(ctx->stream));
        return HSPS_OK;
    });
}
int hsps_memcpy_d2d(hsps_ctx ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        if (bytes == 0) return HSPS_OK;
        ctx->set_current();
        HSPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return HSPS_OK;
    });
}
int hsps_memset(hsps_ctx ctx, void* dst, int value, size_t bytes) {
    return guarded([&] {
        if (!ctx) throw Error(HSPS_ERR_INVALID, "null context");
        if (bytes == 0) return HSPS_OK;
        ctx->set_current();
        HSPS_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        return HSPS_OK;
    });
}

}  // extern "C"

// Achievable HBM streaming rate: pn = 0.5 r + 0.25 p over FP32 vectors, two
// reads and one write per value in 16-byte nontemporal accesses, four groups
// per lane in flight (the fastest of the read-2 / write-1 forms measured on
// MI355X beyond the Infinity Cache).
constexpr int kProbeBlock = 256, kProbeU = 4, kProbeTile = kProbeBlock * kProbeU * 4;  // values per workgroup
__global__ void __launch_bounds__(kProbeBlock) k_stream_probe(const float* __restrict__ r, const float* __restrict__ p,
                                                              float* __restrict__ pn) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const long long base = (long long)blockIdx.x * kProbeTile;
    f4 rv[kProbeU], pv[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long g = base + (threadIdx.x + u * kProbeBlock) * 4;
        rv[u] = __builtin_nontemporal_load(reinterpret_cast<const f4*>(r + g));
        pv[u] = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p + g));
    }
#pragma unroll
    for (int u = 0; u < kProbeU; ++u) {
        const long long g = base + (threadIdx.x + u * kProbeBlock) * 4;
        __builtin_nontemporal_store(0.5f * rv[u] + 0.25f * pv[u], reinterpret_cast<f4*>(pn + g));
    }
}

extern "C" {

int hsps_stream_probe(hsps_ctx ctx, long long n, int reps, double* gbs) {
    return guarded([&] {
        if (!ctx || !gbs || reps < 1) throw Error(HSPS_ERR_INVALID, "hsps_stream_probe: bad argument");
        const long long tiles = n / kProbeTile;
        if (tiles < 1 || tiles > (1LL << 30)) throw Error(HSPS_ERR_INVALID, "hsps_stream_probe: n out of range");
        const size_t bytes = (size_t)tiles * kProbeTile * sizeof(float);
        ctx->set_current();
        hipStream_t s = ctx->stream;
        DevBuf r(ctx, bytes), p(ctx, bytes), pn(ctx, bytes);
        HSPS_HIP(hipMemsetAsync(r.p, 0, bytes, s));
        HSPS_HIP(hipMemsetAsync(p.p, 0, bytes, s));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        std::vector<float> ms;
        try {
            HSPS_HIP(hipEventCreate(&e0));
            HSPS_HIP(hipEventCreate(&e1));
            for (int k = -3; k < reps; ++k) {  // three untimed launches first
                HSPS_HIP(hipEventRecord(e0, s));
                hipLaunchKernelGGL(k_stream_probe, dim3((unsigned)tiles), dim3(kProbeBlock), 0, s, r.as<float>(),
                                   p.as<float>(), pn.as<float>());
                HSPS_CHECK_LAUNCH();
                HSPS_HIP(hipEventRecord(e1, s));
                HSPS_HIP(hipEventSynchronize(e1));
                float t = 0.0f;
                HSPS_HIP(hipEventElapsedTime(&t, e0, e1));
                if (k >= 0) ms.push_back(t);
            }
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::sort(ms.begin(), ms.end());
        *gbs = 3.0 * (double)bytes / (ms[ms.size() / 2] * 1e-3) * 1e-9;
        return HSPS_OK;
    });
}

int hsps_csr_create(hsps_ctx ctx, int rows, int cols, int nnz, const int* rp, const int* ci, const double* v,
                    hsps_csr* out) {
    return guarded([&] {
        if (!ctx || !out || !rp || (nnz > 0 && (!ci || !v)))
            throw Error(HSPS_ERR_INVALID, "hsps_csr_create: null argument");
        validate_csr(rows, cols, nnz, rp, ci);
        auto A = make_csr(ctx, rows, cols, std::vector<int>(rp, rp + rows + 1), std::vector<int>(ci, ci + nnz),
                          std::vector<double>(v, v + nnz));
        auto* h = new hsps_csr_s;
        static_cast<Csr&>(*h) = std::move(*A);
        ctx_retain(ctx);
        *out = h;
        return HSPS_OK;
    });
}

int hsps_csr_destroy(hsps_csr A) {
    return guarded([&] {
        if (A) csr_release(A);
        return HSPS_OK;
    });
}

int hsps_csr_info(hsps_csr A, int* rows, int* cols, int* nnz) {
    return guarded([&] {
        if (!A) throw Error(HSPS_ERR_INVALID, "null matrix");
        if (rows) *rows = A->rows;
        if (cols) *cols = A->cols;
        if (nnz) *nnz = A->nnz;
        return HSPS_OK;
    });
}

int hsps_csr_update_values(hsps_csr A, const double* v) {
    return guarded([&] {
        if (!A || (!v && A->nnz)) throw Error(HSPS_ERR_INVALID, "null argument");
        HSPS_HIP(hipStreamSynchronize(A->ctx->stream));
        std::copy(v, v + A->nnz, A->h_v.begin());
        A->upload_values();
        A->values_version += 1;
        return HSPS_OK;
    });
}

int hsps_spmv_f64(hsps_csr A, const double* x, double* y) {
    return guarded([&] {
        if (!A || (A->rows && !y) || (A->cols && !x)) throw Error(HSPS_ERR_INVALID, "hsps_spmv_f64: null argument");
        A->ctx->set_current();
        spmv_f64(*A, x, y, A->ctx->stream);
        return HSPS_OK;
    });
}
int hsps_spmv_f32(hsps_csr A, const float* x, float* y) {
    return guarded([&] {
        if (!A || (A->rows && !y) || (A->cols && !x)) throw Error(HSPS_ERR_INVALID, "hsps_spmv_f32: null argument");
        if (!A->v32.p && A->nnz) throw Error(HSPS_ERR_UNSUPPORTED, "matrix has no FP32 copy");
        A->ctx->set_current();
        spmv_f32(*A, x, y, A->ctx->stream);
        return HSPS_OK;
    });
}
int hsps_residual_f64(hsps_csr A, const double* b, const double* x, double* r) {
    return guarded([&] {
        if (!A || !b || !x || !r) throw Error(HSPS_ERR_INVALID, "hsps_residual_f64: null argument");
        if (A->rows != A->cols) throw Error(HSPS_ERR_SIZE, "residual needs a square matrix");
        A->ctx->set_current();
        residual_f64(*A, b, x, r, A->ctx->stream);
        return HSPS_OK;
    });
}

static double ctx_dot(hsps_ctx ctx, int n, const double* x, const double* y) {
    ctx->set_current();
    dot_f64_dev(n, x, y, ctx->red_out.as<double>(), ctx->red_partials.as<double>(), ctx->stream);
    HSPS_HIP(hipMemcpyAsync(ctx->pinned, ctx->red_out.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HSPS_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->pinned[0];
}

int hsps_dot_f64(hsps_ctx ctx, int n, const double* x, const double* y, double* out) {
    return guarded([&] {
        if (!ctx || !out || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_dot_f64: bad argument");
        *out = n == 0 ? 0.0 : ctx_dot(ctx, n, x, y);
        return HSPS_OK;
    });
}
int hsps_norm2_f64(hsps_ctx ctx, int n, const double* x, double* out) {
    return guarded([&] {
        if (!ctx || !out || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_norm2_f64: bad argument");
        *out = n == 0 ? 0.0 : std::sqrt(ctx_dot(ctx, n, x, x));
        return HSPS_OK;
    });
}
int hsps_axpy_f64(hsps_ctx ctx, int n, double a, const double* x, double* y) {
    return guarded([&] {
        if (!ctx || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_axpy_f64: bad argument");
        if (n) { ctx->set_current(); axpy_f64(n, a, x, y, ctx->stream); }
        return HSPS_OK;
    });
}
int hsps_axpby_f64(hsps_ctx ctx, int n, double a, const double* x, double b, double* y) {
    return guarded([&] {
        if (!ctx || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_axpby_f64: bad argument");
        if (n) { ctx->set_current(); axpby_f64(n, a, x, b, y, ctx->stream); }
        return HSPS_OK;
    });
}
int hsps_scale_f64(hsps_ctx ctx, int n, double a, double* x) {
    return guarded([&] {
        if (!ctx || n < 0) throw Error(HSPS_ERR_INVALID, "hsps_scale_f64: bad argument");
        if (n) { ctx->set_current(); scale_f64(n, a, x, ctx->stream); }
        return HSPS_OK;
    });
}

}  // extern "C"
