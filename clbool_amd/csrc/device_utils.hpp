// device_utils.hpp -- wave64 / workgroup building blocks for CDNA4.
//
// Scans and reductions are written for 64-wide wavefronts: a wave scan is
// six shuffle steps, a workgroup scan adds one LDS exchange of the per-wave
// totals.  Nothing here assumes 32-wide warps.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace clb {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane() { return static_cast<int>(__lane_id()); }
__device__ __forceinline__ int wave_id() { return static_cast<int>(threadIdx.x) / WAVE; }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int l = lane();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T u = __shfl_up(v, o, WAVE);
        if (l >= o) v += u;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, WAVE);
        v = u < v ? u : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, WAVE);
        v = u > v ? u : v;
    }
    return v;
}

// Exclusive scan across a workgroup of WG threads.  `lds` needs WG/64 + 1
// elements.  Returns the thread's exclusive prefix; `total` gets the sum.
template <typename T, int WG>
__device__ __forceinline__ T block_excl_scan(T v, T *lds, T &total) {
    constexpr int NW = WG / WAVE;
    const int l = lane(), w = wave_id();
    T inc = wave_incl_scan(v);
    if (NW == 1) {
        total = __shfl(inc, WAVE - 1, WAVE);
        return inc - v;
    }
    if (l == WAVE - 1) lds[w] = inc;
    __syncthreads();
    if (w == 0) {
        T x = l < NW ? lds[l] : T(0);
        T xi = wave_incl_scan(x);
        if (l < NW) lds[l] = xi - x;
        if (l == NW - 1) lds[NW] = xi;
    }
    __syncthreads();
    T r = inc - v + lds[w];
    total = lds[NW];
    __syncthreads();
    return r;
}

template <typename T, int WG>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    constexpr int NW = WG / WAVE;
    T s = wave_sum(v);
    if (NW == 1) return s;
    const int l = lane(), w = wave_id();
    if (l == 0) lds[w] = s;
    __syncthreads();
    T t = T(0);
#pragma unroll
    for (int i = 0; i < NW; ++i) t += lds[i];
    __syncthreads();
    return t;
}

// Multiplicative (Fibonacci) hash of a column index into a table of 2^bits
// slots: the high bits of col * 2^32/phi spread consecutive columns apart.
__device__ __forceinline__ uint32_t hash_slot(uint32_t col, int bits) {
    return (col * 0x9E3779B1u) >> (32 - bits);
}

// Ascending bitonic sort of n2 (a power of two) uint32 keys in LDS by a
// workgroup of WG threads.
template <int WG>
__device__ void lds_bitonic_sort(uint32_t *keys, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < (n2 >> 1); t += WG) {
                const uint32_t lo = 2 * j * (t / j) + (t % j);
                const uint32_t hi = lo + j;
                const bool up = (lo & k) == 0;
                uint32_t a = keys[lo], b = keys[hi];
                if ((a > b) == up) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Ascending bitonic sort of one value per lane across a wave (64 values).
__device__ __forceinline__ uint32_t wave_sort64(uint32_t v) {
    const int l = lane();
#pragma unroll
    for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t x = __shfl_xor(v, j, WAVE);
            const bool up = (l & k) == 0;
            const bool lower = (l & j) == 0;
            const uint32_t mn = x < v ? x : v, mx = x < v ? v : x;
            v = (lower == up) ? mn : mx;
        }
    }
    return v;
}

// Ascending sort of a[0, n) in LDS by a workgroup of WG threads: waves sort
// 64-element blocks in registers (the tail padded with 0xFFFFFFFF), then
// log2(n/64) rounds of merge-path merges ping-pong between a and b (scratch
// of at least n rounded up to 64 words, disjoint from a).  Each thread
// produces K consecutive outputs of a round, K a power of two <= 64, so no
// thread's outputs straddle two merged pairs.  Returns the buffer holding
// the result, padded with 0xFFFFFFFF up to the next multiple of 64.
template <int WG>
__device__ uint32_t *lds_merge_sort(uint32_t *a, uint32_t *b, uint32_t n) {
    constexpr int NW = WG / WAVE;
    const int l = lane();
    const uint32_t n64 = (n + 63u) & ~63u;
    for (uint32_t blk = (uint32_t)wave_id(); blk < n64 / 64u; blk += NW) {
        const uint32_t t = blk * 64u + (uint32_t)l;
        const uint32_t v = wave_sort64(t < n ? a[t] : 0xFFFFFFFFu);
        a[t] = v;
    }
    __syncthreads();
    uint32_t per = n64 / WG;
    uint32_t K = 1;
    while (K < per && K < 64u) K <<= 1;
    uint32_t *src = a, *dst = b;
    for (uint32_t w = 64; w < n64; w <<= 1) {
        for (uint32_t o0 = threadIdx.x * K; o0 < n64; o0 += WG * K) {
            const uint32_t s = o0 & ~(2u * w - 1u);
            const uint32_t la = n64 - s < w ? n64 - s : w;
            const uint32_t rest = n64 - s - la;
            const uint32_t lb = rest < w ? rest : w;
            const uint32_t *A = src + s, *B = src + s + la;
            const uint32_t d = o0 - s;
            uint32_t lo = d > lb ? d - lb : 0u, hi = d < la ? d : la;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (A[mid] <= B[d - 1u - mid]) lo = mid + 1u;
                else hi = mid;
            }
            uint32_t i = lo, j = d - lo;
            uint32_t av = i < la ? A[i] : 0xFFFFFFFFu, bv = j < lb ? B[j] : 0xFFFFFFFFu;
            for (uint32_t k = 0; k < K; ++k) {
                const bool takeA = j >= lb || (i < la && av <= bv);
                dst[o0 + k] = takeA ? av : bv;
                if (takeA) {
                    ++i;
                    av = i < la ? A[i] : 0xFFFFFFFFu;
                } else {
                    ++j;
                    bv = j < lb ? B[j] : 0xFFFFFFFFu;
                }
            }
        }
        __syncthreads();
        uint32_t *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

}  // namespace dev
}  // namespace clb
