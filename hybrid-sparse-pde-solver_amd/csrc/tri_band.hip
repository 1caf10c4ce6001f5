// =============================================================================
// tri_band.hip — wavefront triangular sweeps for banded factors (layout and
// schedule in tri_band.hpp).
// =============================================================================
#include "tri_band.hpp"

#include <algorithm>
#include <cstdlib>

#include "device_common.hpp"

namespace hsps_dev {

constexpr unsigned char kBandNone = 0xFF;  // empty entry slot
constexpr int kBandWaves = 16;             // most waves of the one workgroup
constexpr int kBandChunk = 16;             // steps between handoff checks (T is a multiple of it)
constexpr int kBandHistMax = 4;            // longest history a sweep may need
constexpr int kBandMinRows = 4096;         // smaller factors stay on the level sweeps
constexpr int kBandLdsBytes = 128 << 10;   // handoff rings: one band's last grid row per wave
constexpr int kBandSpinMax = 1 << 22;      // waits give up (and flag an error) after this many polls
constexpr int kBandTail = 16 * 64;         // zero positions after the last band (the deepest prefetch)

template <int K>
struct BandCode;
template <>
struct BandCode<2> {
    using type = unsigned short;
};
template <>
struct BandCode<4> {
    using type = unsigned int;
};
template <>
struct BandCode<8> {
    using type = unsigned long long;
};

template <int K, class V, class T>
struct BandRow {
    V v[K];
    typename BandCode<K>::type c;
    T x, dv;
};

// Lane l gets lane l-1's v; lane 0 gets `first` (DPP wave_shr:1, bound_ctrl
// off: a lane without a source keeps the old value).
__device__ __forceinline__ float shift_up(float v, float first) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double shift_up(double v, double first) {
    const long long b = __double_as_longlong(v), f = __double_as_longlong(first);
    const int lo = __builtin_amdgcn_update_dpp((int)(f & 0xffffffffLL), (int)(b & 0xffffffffLL), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(f >> 32), (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T>
__device__ __forceinline__ T lane_value(T v, int l) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

// One workgroup of 64 NW threads; wave w takes bands w, w + NW, ...  Lane l
// of band `band` owns grid row 64 band + l and solves column tau - S l at
// step tau.  Histories: own[m] = this lane's result m+1 steps ago, prv[m] =
// lane l-1's result m+1 steps ago (for lane 0: the previous band's last grid
// row at column tau + S - 1 - m).  Entry code m < H selects own[m], H <= m <
// 2H selects prv[m - H].
//
// Handoff: lane 63 writes its grid row into the wave's LDS ring (W slots,
// one band) as it goes; the next band's wave reads it there (a broadcast
// read).  Every kBandChunk steps a wave publishes how far its producer and
// consumer sides got (pub, con: pair-stream round x W + columns) and checks,
// before the next chunk, that the columns it will read are written (pub of
// the previous wave) and that the ring slots it will overwrite were read by
// its consumer one round ago (con of the next wave).  Both waits point to an
// earlier band (band - 1, band - NW + 1), so the earliest unfinished band
// never waits: the sweep cannot deadlock.
template <int K, int H, class V, class T, bool UPPER>
__global__ void __launch_bounds__(64 * kBandWaves)
k_tri_band(int NB, int T_band, int W, int S, const V* __restrict__ val, const typename BandCode<K>::type* __restrict__ code,
           const T* __restrict__ invd, const T* __restrict__ in, T* __restrict__ out, int* err, const int* skip) {
    static_assert(H >= 1 && H <= kBandHistMax, "history depth");
    constexpr int P = kBandChunk;
    constexpr int kAhead = (K <= 2 ? 8 : K <= 4 ? 4 : 2) * (sizeof(T) == 4 ? 2 : 1);
    static_assert(P % kAhead == 0, "ring rounds fit a chunk");
    extern __shared__ __attribute__((aligned(16))) unsigned char band_lds[];
    T* ring = (T*)band_lds;  // [NW][W]
    __shared__ int pub[kBandWaves], con[kBandWaves], stop;
    if (skip && *skip) return;
    const int NW = (int)(blockDim.x >> 6), w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    if (threadIdx.x < kBandWaves) pub[threadIdx.x] = con[threadIdx.x] = 0;
    if (threadIdx.x == 0) stop = 0;
    __syncthreads();
    const int wn = (w + 1) % NW, wp = (w + NW - 1) % NW;  // consumer and producer waves
    T* my_ring = ring + (size_t)w * W;
    const T* prev_ring = ring + (size_t)wp * W;
    int alive = 1;  // wave-uniform
    // polls `ctr` until it reaches `need` (uniform); 0 when the sweep gives up
    auto wait_for = [&](const int* ctr, int need) {
        for (int it = 0;; ++it) {
            const int v = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (v >= need) break;
            const int st = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (st || it > kBandSpinMax) {
                if (!st && lane == 0) {
                    __hip_atomic_store(&stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    *err = 1;
                }
                alive = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };
    for (int band = w; band < NB && alive; band += NW) {
        const int round = band / NW;
        const bool has_prev = band > 0;
        // pair-stream rounds: this band's output is round `round` of (w -> wn);
        // its input is round kin of (wp -> w)
        const int kin = w == 0 ? round - 1 : round;
        T own[H], prv[H];
#pragma unroll
        for (int m = 0; m < H; ++m) own[m] = prv[m] = T(0);
        const long long base = (long long)band * T_band * 64 + lane;
        const V* vp = val + base * K;
        const typename BandCode<K>::type* cp = code + base;
        const T* xp = in + base;
        const T* dp = UPPER ? invd + base : nullptr;
        T* op = out + base;
        auto push = [&](T res, T hv) {
            const T p = shift_up(res, hv);
#pragma unroll
            for (int m = H - 1; m > 0; --m) {
                own[m] = own[m - 1];
                prv[m] = prv[m - 1];
            }
            own[0] = res;
            prv[0] = p;
        };
        auto handoff = [&](int col) -> T {  // the previous band's last grid row at `col` (0 past the row)
            return (has_prev && col < W) ? prev_ring[col] : T(0);
        };
        // warm-up: lane 0's history takes columns 0 .. S-1 of the previous band
        if (has_prev) wait_for(&pub[wp], kin * W + min(W, S));
        for (int tau = -S; tau < 0; ++tau) push(T(0), handoff(tau + S));
        BandRow<K, V, T> rr[kAhead];
        auto load = [&](BandRow<K, V, T>& R, int tau) {  // rows past the band: the next band's or the zero tail
            const long long q = (long long)tau * 64;
#pragma unroll
            for (int k = 0; k < K; ++k) R.v[k] = vp[q * K + k];
            R.c = cp[q];
            R.x = xp[q];
            if (UPPER) R.dv = dp[q];
        };
#pragma unroll
        for (int q = 0; q < kAhead; ++q) load(rr[q], q);
        for (int c0 = 0; c0 < T_band && alive; c0 += P) {
            // columns this chunk reads (tau + S) and writes (tau - 63 S)
            if (has_prev) wait_for(&pub[wp], kin * W + min(W, c0 + P + S));
            const int wr_end = min(W, c0 + P - 63 * S);
            if (round > 0 && wr_end > 0) wait_for(&con[wn], (round - 1) * W + wr_end);
            if (!alive) break;
            for (int q0 = 0; q0 < P; q0 += kAhead) {
#pragma unroll
                for (int q = 0; q < kAhead; ++q) {
                    const int tau = c0 + q0 + q;
                    BandRow<K, V, T>& R = rr[q];
                    T s = R.x;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const unsigned m = (unsigned)((R.c >> (8 * k)) & 0xFF);
                        T h = own[0];
#pragma unroll
                        for (int mm = 1; mm < 2 * H; ++mm) h = m == (unsigned)mm ? (mm < H ? own[mm] : prv[mm - H]) : h;
                        const T t = s - (T)R.v[k] * h;
                        s = m != kBandNone ? t : s;
                    }
                    if (UPPER) s = s * R.dv;
                    op[(long long)tau * 64] = s;
                    const int c63 = tau - 63 * S;
                    if (c63 >= 0 && c63 < W && lane == 63) my_ring[c63] = s;
                    push(s, handoff(tau + S));
                    load(R, tau + kAhead);
                }
            }
            // publish: LDS ring writes before the counters
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            if (lane == 0) {
                __hip_atomic_store(&pub[w], round * W + max(0, wr_end), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (has_prev)
                    __hip_atomic_store(&con[w], kin * W + min(W, c0 + P + S), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (lane == 0) {  // the band is done: all of it written and read
            __hip_atomic_store(&pub[w], round * W + W, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (has_prev) __hip_atomic_store(&con[w], kin * W + W, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

template <class T>
__global__ void k_band_gather(int n, const int* __restrict__ sk, const T* __restrict__ in, T* __restrict__ in_sk,
                              const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) in_sk[sk[i]] = in[i];
}
template <class T>
__global__ void k_band_scatter(int n, const int* __restrict__ sk, const T* __restrict__ out_sk, T* __restrict__ out,
                               const int* skip) {
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = out_sk[sk[i]];
}

// ---- host: layout ---------------------------------------------------------------------------
TriBand make_tri_band(Ctx* c, const hsps_host::HostCsr& F, bool lower, const std::vector<double>& invd) {
    TriBand B;
    const char* env = std::getenv("HSPS_ILU_BAND");  // 0: level sweeps only (A/B measurements)
    if (env && std::atoi(env) == 0) return B;
    const int n = F.rows;
    if (n < kBandMinRows) return B;
    auto is_dep = [&](int i, int j) { return lower ? j < i : j > i; };
    // pass 1: width, entries per row
    int W = 0, K = 0;
    for (int i = 0; i < n; ++i) {
        const int ip = lower ? i : n - 1 - i;
        int cnt = 0;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {
            const int j = F.ci[k];
            if (!is_dep(i, j)) continue;
            const int cp = lower ? j : n - 1 - j;
            W = std::max(W, ip - cp);
            ++cnt;
        }
        K = std::max(K, cnt);
    }
    if (W == 0 || K > 8 || W > 65535) return B;  // no dependencies (one level), or too wide
    // pass 2: skew and own look-back
    int S = 1, age_max = 0;
    for (int i = 0; i < n; ++i) {
        const int ip = lower ? i : n - 1 - i, jj = ip % W;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {
            const int j = F.ci[k];
            if (!is_dep(i, j)) continue;
            const int d = ip - (lower ? j : n - 1 - j);
            if (d <= jj) age_max = std::max(age_max, d);
            else S = std::max(S, W - d + 1);
        }
    }
    if (age_max > kBandHistMax || S > kBandHistMax) return B;
    const int Kt = (K <= 2 && age_max <= 1 && S == 1) ? 2 : K <= 4 ? 4 : 8;
    const int Hh = Kt == 2 ? 1 : kBandHistMax;
    const long long G = (n + (long long)W - 1) / W, NB = (G + 63) / 64;
    const long long T = ((long long)W + 63LL * S + kBandChunk - 1) / kBandChunk * kBandChunk;
    const long long n_sk = NB * T * 64;
    if (n_sk > 5LL * n / 2 || n_sk >= (1LL << 31)) return B;  // padding too large
    B.K = Kt;
    B.H = Hh;
    B.W = W;
    B.S = S;
    B.NB = (int)NB;
    B.T = (int)T;
    // waves: at most 16, one per band, and their handoff rings (W values each) within the LDS budget
    B.NW64 = (int)std::min<long long>(std::min<long long>(kBandWaves, NB), kBandLdsBytes / (8LL * W));
    B.NW32 = (int)std::min<long long>(std::min<long long>(kBandWaves, NB), kBandLdsBytes / (4LL * W));
    if (B.NW64 < 1) return TriBand{};
    B.n_sk = n_sk;
    // the sweep's prefetch reads up to kBandTail positions past the last band: zero rows
    const size_t n_all = (size_t)n_sk + kBandTail;
    std::vector<double> v(n_all * Kt, 0.0), dv(lower ? 0 : n_all, 0.0);
    std::vector<unsigned char> code(n_all * Kt, kBandNone);
    std::vector<int> sk(n);
    for (int i = 0; i < n; ++i) {
        const int ip = lower ? i : n - 1 - i;
        const long long g = ip / W, jj = ip % W, band = g / 64, l = g % 64;
        const long long p = (band * T + jj + (long long)S * l) * 64 + l;
        sk[i] = (int)p;
        int e = 0;
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {  // ascending columns: the level sweeps' order
            const int j = F.ci[k];
            if (!is_dep(i, j)) continue;
            const int d = ip - (lower ? j : n - 1 - j);
            const int m = d <= jj ? d - 1 : Hh + (S - 1 - (W - d));
            v[(size_t)p * Kt + e] = F.v[k];
            code[(size_t)p * Kt + e] = (unsigned char)m;
            ++e;
        }
        if (!lower) dv[(size_t)p] = invd[i];
    }
    auto up = [&](DevBuf& d, const void* h, size_t bytes) {
        d = DevBuf(c, std::max<size_t>(bytes, 8));
        upload_on(c->stream, d.p, h, bytes);
    };
    up(B.val64, v.data(), v.size() * sizeof(double));
    {
        std::vector<float> f(v.begin(), v.end());
        up(B.val32, f.data(), f.size() * sizeof(float));
    }
    up(B.code, code.data(), code.size());
    if (!lower) {
        up(B.invd64, dv.data(), dv.size() * sizeof(double));
        std::vector<float> f(dv.begin(), dv.end());
        up(B.invd32, f.data(), f.size() * sizeof(float));
    }
    up(B.sk_of_row, sk.data(), sk.size() * sizeof(int));
    B.in64 = DevBuf(c, n_all * sizeof(double));
    B.in32 = DevBuf(c, n_all * sizeof(float));
    B.out64 = DevBuf(c, n_all * sizeof(double));
    B.out32 = DevBuf(c, n_all * sizeof(float));
    fill_on(c->stream, B.in64.p, 0, n_all * sizeof(double));
    fill_on(c->stream, B.in32.p, 0, n_all * sizeof(float));
    B.ok = 1;
    return B;
}

// ---- host: sweep ------------------------------------------------------------------------------
template <int K, int H, class T, bool UPPER>
static void band_launch(const TriBand& B, const T* in_sk, T* out_sk, int* err, hipStream_t s, const int* skip) {
    using Code = typename BandCode<K>::type;
    const Code* code = B.code.as<Code>();
    constexpr bool dbl = std::is_same<T, double>::value;
    const int NW = dbl ? B.NW64 : B.NW32;
    const size_t lds = sizeof(T) * (size_t)NW * B.W;
    auto kern = k_tri_band<K, H, T, T, UPPER>;
    static const bool lds_ok = [&] {  // once per instantiation: allow the full handoff ring budget
        return hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, kBandLdsBytes) ==
               hipSuccess;
    }();
    if (!lds_ok) throw Error(HSPS_ERR_HIP, "ILU band sweep: cannot reserve the LDS handoff rings");
    hipLaunchKernelGGL(kern, dim3(1), dim3(64 * NW), lds, s, B.NB, B.T, B.W, B.S,
                       dbl ? (const T*)B.val64.p : (const T*)B.val32.p, code,
                       UPPER ? (dbl ? (const T*)B.invd64.p : (const T*)B.invd32.p) : nullptr, in_sk, out_sk, err,
                       skip);
}

template <class T>
void tri_band_sweep(const TriBand& B, bool upper, int n, const T* in, T* out, int* err, hipStream_t s,
                    const int* skip) {
    constexpr bool dbl = std::is_same<T, double>::value;
    T* in_sk = dbl ? (T*)B.in64.p : (T*)B.in32.p;
    T* out_sk = dbl ? (T*)B.out64.p : (T*)B.out32.p;
    const int* sk = B.sk_of_row.as<int>();
    hipLaunchKernelGGL(k_band_gather<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sk, in, in_sk, skip);
    auto go = [&](auto up) {
        constexpr bool U = decltype(up)::value;
        if (B.K == 2) band_launch<2, 1, T, U>(B, in_sk, out_sk, err, s, skip);
        else if (B.K == 4) band_launch<4, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
        else band_launch<8, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
    };
    if (upper) go(std::true_type{});
    else go(std::false_type{});
    hipLaunchKernelGGL(k_band_scatter<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sk, (const T*)out_sk, out, skip);
    HSPS_CHECK_LAUNCH();
}

template void tri_band_sweep<double>(const TriBand&, bool, int, const double*, double*, int*, hipStream_t,
                                     const int*);
template void tri_band_sweep<float>(const TriBand&, bool, int, const float*, float*, int*, hipStream_t, const int*);

}  // namespace hsps_dev
