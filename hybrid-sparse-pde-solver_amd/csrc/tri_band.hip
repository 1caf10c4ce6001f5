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
constexpr int kBandChunk = 16;             // one-workgroup sweep: steps between handoff checks
constexpr int kBandMcChunk = 32;           // multi-workgroup sweep: steps between handoff checks
                                           // (C3 mixed: 32 -> 14.8 ms, 16 -> 15.5 ms)
constexpr int kBandTRound = 32;            // T (steps per band) is a multiple of both chunks
constexpr int kBandHistMax = 4;            // longest history a sweep may need
constexpr int kBandMinRows = 4096;         // smaller factors stay on the level sweeps
constexpr int kBandLdsBytes = 128 << 10;   // handoff rings: one band's last grid row per wave
constexpr int kBandSpinMax = 1 << 22;      // waits give up (and flag an error) after this many polls
constexpr int kBandTail = 32 * 64;         // zero positions after the last band (the deepest prefetch)

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

// ---- several workgroups -------------------------------------------------------------------
// The same sweep with its waves spread over workgroups of kBandMcWaves (one
// wave per SIMD), so each wave issues at the rate of a whole SIMD instead of
// a quarter of it.  Bands hand their last grid row to the next band's wave
// through global memory: per pair of consecutive waves, two rounds of W
// slots, each a 64-bit word (tag << 32 | 32 value bits; FP64 values take two
// words) written and read with relaxed device-scope atomics, so a slot is
// valid exactly when its tag names the round being read: no flags, no
// fences, nothing to order.  A producer overwrites a slot of round r - 2
// only after its consumer's progress counter shows that round read.  Waits
// point to earlier bands only (band - 1, band - 2 x waves + 1), so the
// earliest unfinished band always proceeds; every wait also gives up after
// kBandWaitCycles of the 100 MHz clock (flagging err) instead of hanging.
constexpr int kBandMcWaves = 4;
constexpr int kBandMcMaxWaves = 64;  // at most one wave per band, 16 workgroups
constexpr long long kBandWaitTicks = 200000000LL;  // 2 s of s_memrealtime

template <class T>
struct SlotWords {
    static constexpr int N = sizeof(T) / 4;  // 64-bit slot words per value
};

__device__ __forceinline__ unsigned long long slot_ld(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void slot_st(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FIVE (two entry slots, one-deep histories, every slot's source fixed: the
// neighbour's result first then this lane's own for the lower sweep, the
// reverse for the upper one, an empty slot holding 0 where its source is a
// +0 padding value): no slot codes at all.
template <int K, int H, class V, class T, bool UPPER, bool FIVE>
__global__ void __launch_bounds__(64 * kBandMcWaves)
k_tri_band_mc(int NB, int T_band, int W, int S, int NWT, const V* __restrict__ val,
              const typename BandCode<K>::type* __restrict__ code, const T* __restrict__ invd,
              const T* __restrict__ in, T* __restrict__ out, unsigned long long* slots, int* con, int* err,
              const int* skip) {
    static_assert(H >= 1 && H <= kBandHistMax, "history depth");
    constexpr int P = kBandMcChunk;
    constexpr int SW = SlotWords<T>::N;
    // steps of loads in flight: one wave per SIMD leaves registers for a deep
    // ring; at most 16 (FP32 five-point factors: 32 steps ahead made the
    // conv-diff 1024^2 mixed FGMRES + ILU(0) solve 14.1 ms against 13.3 ms at
    // 16; 8 is slower again)
    constexpr int kAhead = std::min(16, (K <= 2 ? 16 : K <= 4 ? 8 : 4) * (sizeof(T) == 4 ? 2 : 1));
    static_assert(P % kAhead == 0, "ring rounds fit a chunk");
    static_assert(P <= 64, "a chunk's handoff fits the lanes");
    if (skip && *skip) return;
    const int gw = (int)blockIdx.x * kBandMcWaves + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (gw >= NWT) return;
    const int lane = (int)(threadIdx.x & 63);
    const int gwp = (gw + NWT - 1) % NWT, gwn = (gw + 1) % NWT;
    int alive = 1;  // wave-uniform
    const long long t_start = (long long)__builtin_amdgcn_s_memrealtime();
    auto timed_out = [&]() {
        if ((long long)__builtin_amdgcn_s_memrealtime() - t_start > kBandWaitTicks) {
            if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            alive = 0;
            return true;
        }
        return false;
    };
    for (int band = gw; band < NB && alive; band += NWT) {
        const int round = band / NWT;
        const bool has_prev = band > 0;
        const int kin = gw == 0 ? round - 1 : round;  // input pair-stream round
        const unsigned tag_in = (unsigned)kin + 1u, tag_out = (unsigned)round + 1u;
        unsigned long long* out_slots = slots + (size_t)(gw * 2 + (round & 1)) * W * SW;
        const unsigned long long* in_slots = slots + (size_t)(gwp * 2 + ((kin + 2) & 1)) * W * SW;
        // the handoff of columns [c, c + P): lane q < P holds column c + q (0 past the row).
        // issue() loads the slot words; take() checks their tags and, while some
        // lane's is not the round's yet, loads again (uniform).
        struct Hand {
            unsigned long long w[SW];
        };
        auto issue = [&](int c, Hand& h) {
            const int col = c + lane;
#pragma unroll
            for (int k = 0; k < SW; ++k)
                h.w[k] = (has_prev && lane < P && col < W) ? slot_ld(in_slots + (size_t)col * SW + k)
                                                           : ((unsigned long long)tag_in << 32);
        };
        auto take = [&](int c, Hand& h) -> T {
            for (;;) {
                bool ok = true;
                unsigned long long bits = 0;
#pragma unroll
                for (int k = 0; k < SW; ++k) {
                    ok = ok && (unsigned)(h.w[k] >> 32) == tag_in;
                    bits |= (h.w[k] & 0xffffffffULL) << (32 * k);
                }
                if (__builtin_amdgcn_ballot_w64(!ok) == 0) {
                    if constexpr (sizeof(T) == 4) return __int_as_float((int)(unsigned)bits);
                    else return __longlong_as_double((long long)bits);
                }
                if (timed_out()) return T(0);
                __builtin_amdgcn_s_sleep(2);
                issue(c, h);
            }
        };
        // waits until the consumer of this wave's slots has read round `round - 2` up to column `cols`
        auto wait_consumer = [&](int cols) {
            const int need = (round - 2) * W + cols;
            while (alive) {
                const int v = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(con + gwn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (v >= need || timed_out()) break;
                __builtin_amdgcn_s_sleep(2);
            }
        };
        T own[H], prv[H];
#pragma unroll
        for (int m = 0; m < H; ++m) own[m] = prv[m] = T(0);
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
        // warm-up: lane 0's history takes columns 0 .. S-1 of the previous band
        Hand hn;
        {
            issue(0, hn);
            const T h0 = take(0, hn);
            for (int tau = -S; tau < 0; ++tau) push(T(0), lane_value(h0, tau + S));
        }
        issue(S, hn);  // the first chunk's handoff
        // uniform band bases, 32-bit lane offsets (a band holds < 2^31 / K positions)
        const long long base = (long long)band * T_band * 64;
        const V* vp = val + base * K;
        const typename BandCode<K>::type* cp = code + base;
        const T* xp = in + base;
        const T* dp = UPPER ? invd + base : nullptr;
        T* op = out + base;
        BandRow<K, V, T> rr[kAhead];
        auto load = [&](BandRow<K, V, T>& R, int tau) {  // rows past the band: the next band's or the zero tail
            const unsigned q = (unsigned)tau * 64u + (unsigned)lane;
#pragma unroll
            for (int k = 0; k < K; ++k) R.v[k] = vp[q * K + k];
            if (!FIVE) R.c = cp[q];
            R.x = xp[q];
            if (UPPER) R.dv = dp[q];
        };
#pragma unroll
        for (int q = 0; q < kAhead; ++q) load(rr[q], q);
        for (int c0 = 0; c0 < T_band && alive; c0 += P) {
            const T hcur = take(c0 + S, hn);  // lane q: column c0 + S + q, pushed at step c0 + q
            issue(c0 + P + S, hn);             // the next chunk's, in flight during this one
            T wr = T(0);                       // lane q: lane 63's result at step c0 + q
            for (int q0 = 0; q0 < P; q0 += kAhead) {
#pragma unroll
                for (int q = 0; q < kAhead; ++q) {
                    const int tau = c0 + q0 + q;
                    BandRow<K, V, T>& R = rr[q];
                    T s = R.x;
                    if constexpr (FIVE) {
                        s = s - (T)R.v[0] * (UPPER ? own[0] : prv[0]);
                        s = s - (T)R.v[1] * (UPPER ? prv[0] : own[0]);
                    } else {
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const unsigned m = (unsigned)((R.c >> (8 * k)) & 0xFF);
                            T h = own[0];
#pragma unroll
                            for (int mm = 1; mm < 2 * H; ++mm)
                                h = m == (unsigned)mm ? (mm < H ? own[mm] : prv[mm - H]) : h;
                            const T t = s - (T)R.v[k] * h;
                            s = m != kBandNone ? t : s;
                        }
                    }
                    if (UPPER) s = s * R.dv;
                    op[(unsigned)tau * 64u + (unsigned)lane] = s;
                    const T s63 = lane_value(s, 63);
                    wr = lane == q0 + q ? s63 : wr;
                    push(s, lane_value(hcur, q0 + q));
                    load(R, tau + kAhead);
                }
            }
            // lane 63's columns of this chunk: c0 - 63 S + q, q < P
            const int wc = c0 - 63 * S + lane;
            const int wr_end = min(W, c0 + P - 63 * S);
            if (wr_end > 0) {
                if (round >= 2) wait_consumer(wr_end);
                if (lane < P && wc >= 0 && wc < W) {
                    unsigned long long bits;
                    if constexpr (sizeof(T) == 4) bits = (unsigned)__float_as_int(wr);
                    else bits = (unsigned long long)__double_as_longlong(wr);
#pragma unroll
                    for (int k = 0; k < SW; ++k)
                        slot_st(out_slots + (size_t)wc * SW + k,
                                ((unsigned long long)tag_out << 32) | ((bits >> (32 * k)) & 0xffffffffULL));
                }
            }
            if (has_prev && lane == 0)
                __hip_atomic_store(con + gw, kin * W + min(W, c0 + P + S), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (has_prev && lane == 0)
            __hip_atomic_store(con + gw, kin * W + W, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Clears the handoff slots (tags) and progress counters a multi-workgroup
// sweep starts from (a kernel, not a memset: the sweeps may be captured).
__global__ void k_band_reset(long long n_words, unsigned long long* slots, int n_con, int* con, const int* skip) {
    if (skip && *skip) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n_words; i += (long long)gridDim.x * blockDim.x)
        slots[i] = 0ULL;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_con; i += gridDim.x * blockDim.x) con[i] = 0;
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
    const long long T = ((long long)W + 63LL * S + kBandTRound - 1) / kBandTRound * kBandTRound;
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
    std::vector<double> v(n_all * Kt, 0.0), dv(lower ? 0 : n_all, 0.0), fv(Kt == 2 ? n_all * 2 : 0, 0.0);
    std::vector<unsigned char> code(n_all * Kt, kBandNone);
    // five-point form (two slots, one-deep histories): the summation order puts the
    // neighbour's entry first for the lower sweep and this lane's own first for
    // the upper one, and a missing entry's source is a +0 padding value (own at
    // column 0, neighbour in grid row 0), so slot k can hold a fixed source
    // with 0 standing in for a missing entry (s - 0 * +0 = s, bit for bit)
    bool five = Kt == 2;
    const int src0 = lower ? Hh : 0;  // code of slot 0's source (slot 1: the other one)
    std::vector<int> sk(n);
    for (int i = 0; i < n; ++i) {
        const int ip = lower ? i : n - 1 - i;
        const long long g = ip / W, jj = ip % W, band = g / 64, l = g % 64;
        const long long p = (band * T + jj + (long long)S * l) * 64 + l;
        sk[i] = (int)p;
        int e = 0, slot_of[2] = {-1, -1};
        double f[2] = {0.0, 0.0};
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {  // ascending columns: the level sweeps' order
            const int j = F.ci[k];
            if (!is_dep(i, j)) continue;
            const int d = ip - (lower ? j : n - 1 - j);
            const int m = d <= jj ? d - 1 : Hh + (S - 1 - (W - d));
            v[(size_t)p * Kt + e] = F.v[k];
            code[(size_t)p * Kt + e] = (unsigned char)m;
            if (e < 2) {
                slot_of[e] = m == src0 ? 0 : 1;
                f[slot_of[e]] = F.v[k];
            }
            ++e;
        }
        if (five) {
            // sources in fixed slot order, each at most once
            if (e == 2 && !(slot_of[0] == 0 && slot_of[1] == 1)) five = false;
            const bool has[2] = {(e >= 1 && slot_of[0] == 0), (e == 2) || (e == 1 && slot_of[0] == 1)};
            // slot 0 / 1 source: lower = neighbour / own, upper = own / neighbour
            const bool own_missing = lower ? !has[1] : !has[0], nb_missing = lower ? !has[0] : !has[1];
            if ((own_missing && jj != 0) || (nb_missing && g != 0)) five = false;
            fv[2 * (size_t)p] = f[0];
            fv[2 * (size_t)p + 1] = f[1];
        }
        if (!lower) dv[(size_t)p] = invd[i];
    }
    B.five = five;
    if (five) {  // values in slot order; codes name the slots' fixed sources (for the one-workgroup sweep)
        v.swap(fv);
        const int src1 = lower ? 0 : Hh;
        for (int i = 0; i < n; ++i) {
            code[2 * (size_t)sk[i]] = (unsigned char)src0;
            code[2 * (size_t)sk[i] + 1] = (unsigned char)src1;
        }
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
    B.NWT = (int)std::min<long long>(kBandMcMaxWaves, NB);
    B.n_slot_words = 2LL * B.NWT * 2 * W;  // two rounds of W slots per wave, up to two words per value
    B.slots = DevBuf(c, (size_t)B.n_slot_words * sizeof(unsigned long long));
    B.con = DevBuf(c, sizeof(int) * (size_t)B.NWT);
    B.mode = (env && std::atoi(env) == 1) ? 1 : 2;
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

template <int K, int H, class T, bool UPPER>
static void band_launch_mc(const TriBand& B, const T* in_sk, T* out_sk, int* err, hipStream_t s, const int* skip) {
    using Code = typename BandCode<K>::type;
    constexpr bool dbl = std::is_same<T, double>::value;
    hipLaunchKernelGGL(k_band_reset, dim3(std::min<long long>(64, (B.n_slot_words + 255) / 256)), dim3(256), 0, s,
                       B.n_slot_words, B.slots.as<unsigned long long>(), B.NWT, B.con.as<int>(), skip);
    auto go = [&](auto five) {
        constexpr bool F = decltype(five)::value && K == 2 && H == 1;
        hipLaunchKernelGGL((k_tri_band_mc<K, H, T, T, UPPER, F>), dim3((B.NWT + kBandMcWaves - 1) / kBandMcWaves),
                           dim3(64 * kBandMcWaves), 0, s, B.NB, B.T, B.W, B.S, B.NWT,
                           dbl ? (const T*)B.val64.p : (const T*)B.val32.p, B.code.as<Code>(),
                           UPPER ? (dbl ? (const T*)B.invd64.p : (const T*)B.invd32.p) : nullptr, in_sk, out_sk,
                           B.slots.as<unsigned long long>(), B.con.as<int>(), err, skip);
    };
    if (B.five) go(std::true_type{});
    else go(std::false_type{});
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
        if (B.mode == 1) {
            if (B.K == 2) band_launch<2, 1, T, U>(B, in_sk, out_sk, err, s, skip);
            else if (B.K == 4) band_launch<4, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
            else band_launch<8, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
        } else {
            if (B.K == 2) band_launch_mc<2, 1, T, U>(B, in_sk, out_sk, err, s, skip);
            else if (B.K == 4) band_launch_mc<4, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
            else band_launch_mc<8, kBandHistMax, T, U>(B, in_sk, out_sk, err, s, skip);
        }
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
