// =============================================================================
// features.hip — matrix features for the adaptive ladder
// (AdaptiveSelector::extract_features, adaptive_selector.cpp:38-220): the
// 30-step Lanczos estimate of the spectrum runs on the device (SpMV + FP64
// reductions); the O(nnz) structural scans run on the host copy of the CSR.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <limits>

#include "device_common.hpp"

namespace hsps_dev {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

__global__ void k_fill(int n, double v, double* __restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

// Aq -= alpha q + beta q_prev ; then ||Aq||^2 (adaptive_selector.cpp:80-83)
__global__ void __launch_bounds__(kBlock)
k_lanczos_orth(int n, double* __restrict__ aq, const double* __restrict__ q, const double* __restrict__ qp,
               const double* alpha, double beta, double* partials) {
    const double a = *alpha;
    double red[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double v = aq[i] - (a * q[i] + beta * qp[i]);
        aq[i] = v;
        red[0] += v * v;
    }
    block_partial<1>(red, partials);
}

// q_prev = q ; q = aq / beta
__global__ void k_lanczos_shift(int n, double* __restrict__ qp, double* __restrict__ q, const double* __restrict__ aq,
                                double inv_beta) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        qp[i] = q[i];
        q[i] = aq[i] * inv_beta;
    }
}

static double get_entry(const Csr& A, int i, int j) {  // sparse_matrix.cpp:38-55
    auto b = A.h_ci.begin() + A.h_rp[i], e = A.h_ci.begin() + A.h_rp[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it == e || *it != j) ? 0.0 : A.h_v[it - A.h_ci.begin()];
}

static void features(const Csr& A, double* out) {
    const int n = A.rows;
    if (n == 0 || A.rows != A.cols) throw Error(HSPS_ERR_SIZE, "features: need a non-empty square matrix");
    double fro = 0.0;
    for (double v : A.h_v) fro += v * v;
    fro = std::sqrt(fro);
    double min_dd = std::numeric_limits<double>::infinity();
    for (int i = 0; i < n; ++i) {
        double d = 0.0, off = 0.0;
        for (int k = A.h_rp[i]; k < A.h_rp[i + 1]; ++k) {
            if (A.h_ci[k] == i) d = std::abs(A.h_v[k]); else off += std::abs(A.h_v[k]);
        }
        const double dd = off > kEps ? d / off : std::numeric_limits<double>::infinity();
        min_dd = std::min(min_dd, dd);
    }
    const double diag_dom = std::isinf(min_dd) ? 1.0 : min_dd;

    // Lanczos on the device (adaptive_selector.cpp:66-126)
    Ctx* ctx = A.ctx;
    hipStream_t s = ctx->stream;
    const int K = std::min(n, 30);
    DevBuf q(ctx, n * 8), qp(ctx, n * 8), aq(ctx, n * 8), scal(ctx, 64);
    const int grid = std::max(reduction_blocks(n), 1);
    DevBuf partials(ctx, sizeof(double) * std::max(grid, 1) * 2);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, n, 1.0 / std::sqrt((double)n), q.as<double>());
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, n, 0.0, qp.as<double>());
    std::vector<double> al, be;
    double bj = 0.0;
    double h[2];
    for (int j = 0; j < K; ++j) {
        spmv_f64(A, q.as<double>(), aq.as<double>(), s);
        dot_f64_dev(n, q.as<double>(), aq.as<double>(), scal.as<double>(), partials.as<double>(), s);
        hipLaunchKernelGGL(k_lanczos_orth, dim3(grid), dim3(kBlock), 0, s, n, aq.as<double>(), q.as<double>(),
                           qp.as<double>(), (const double*)scal.as<double>(), bj, partials.as<double>());
        OpStore<1> op;
        op.out = scal.as<double>() + 1;
        launch_finish<1>(partials.as<double>(), grid, op, s);
        HSPS_CHECK_LAUNCH();
        HSPS_HIP(hipMemcpyAsync(h, scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        HSPS_HIP(hipStreamSynchronize(s));
        al.push_back(h[0]);
        bj = std::sqrt(h[1]);
        if (bj < kEps * 1e6) break;
        be.push_back(bj);
        hipLaunchKernelGGL(k_lanczos_shift, dim3(grid_for(n)), dim3(kBlock), 0, s, n, qp.as<double>(), q.as<double>(),
                           (const double*)aq.as<double>(), 1.0 / bj);
        HSPS_CHECK_LAUNCH();
    }
    const int sz = (int)al.size();
    double tmin = al[0], tmax = al[0];
    for (int i = 0; i < sz; ++i) {
        double r = 0.0;
        if (i > 0) r += std::abs(be[i - 1]);
        if (i < sz - 1) r += std::abs(be[i]);
        tmin = std::min(tmin, al[i] - r);
        tmax = std::max(tmax, al[i] + r);
    }
    const double lmax = std::max(tmax, kEps);
    double lmin = tmin;
    if (lmin <= kEps) {  // :99-115
        double mpd = std::numeric_limits<double>::infinity();
        for (int i = 0; i < n; ++i)
            for (int k = A.h_rp[i]; k < A.h_rp[i + 1]; ++k)
                if (A.h_ci[k] == i && A.h_v[k] > kEps) { mpd = std::min(mpd, A.h_v[k]); break; }
        if (std::isinf(mpd)) mpd = kEps;
        lmin = mpd * 0.5;
    }
    lmin = std::max(lmin, kEps);
    // off-diagonal symmetry over the first 200 rows (:167-178)
    double mx = 0.0;
    for (int i = 0; i < std::min(n, 200); ++i)
        for (int k = A.h_rp[i]; k < A.h_rp[i + 1]; ++k) {
            const int j = A.h_ci[k];
            if (j != i) mx = std::max(mx, std::abs(A.h_v[k] - get_entry(A, j, i)));
        }
    const double symres = fro > kEps ? mx / fro : 0.0;
    // SPD sample (:200-219)
    bool sym = true;
    const int sr = std::min(200, n);
    const int stride = std::max(1, n / sr);
    for (int i = 0; i < n && sym; i += stride)
        for (int k = A.h_rp[i]; k < A.h_rp[i + 1] && sym; ++k) {
            const int j = A.h_ci[k];
            if (j == i) continue;
            if (std::abs(A.h_v[k] - get_entry(A, j, i)) > 1e-6 * (1.0 + std::abs(A.h_v[k]))) sym = false;
        }
    bool pd = true;
    for (int i = 0; i < n && pd; ++i)
        for (int k = A.h_rp[i]; k < A.h_rp[i + 1]; ++k)
            if (A.h_ci[k] == i) { if (A.h_v[k] <= 0.0) pd = false; break; }
    out[0] = n;
    out[1] = A.nnz;
    out[2] = diag_dom;
    out[3] = symres;
    out[4] = fro;
    out[5] = lmax / lmin;
    out[6] = lmax;
    out[7] = (sym && pd) ? 1.0 : 0.0;
}

}  // namespace hsps_dev

using namespace hsps_dev;

extern "C" int hsps_features(hsps_csr A, double* out) {
    return guarded([&] {
        if (!A || !out) throw Error(HSPS_ERR_INVALID, "hsps_features: null argument");
        A->ctx->set_current();
        features(*A, out);
        return HSPS_OK;
    });
}
