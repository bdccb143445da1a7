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

}  // namespace dev
}  // namespace clb
