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

// DPP move within rows of 16 lanes (GFX9 encoding: row_shl:n = 0x100 + n,
// row_shr:n = 0x110 + n); lanes whose source lies outside their row get `old`.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_row(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, 0xF, false);
}

// Inclusive sum across a wave with DPP moves only (no LDS traffic): shifts
// 1, 2, 4, 8 within rows of 16 lanes, then row_bcast:15 (0x142, into rows 1
// and 3) and row_bcast:31 (0x143, into rows 2 and 3).  All 64 lanes must be
// active.
__device__ __forceinline__ uint32_t wave_incl_sum_dpp(uint32_t v) {
    v += dpp_row<0x111>(0u, v);
    v += dpp_row<0x112>(0u, v);
    v += dpp_row<0x114>(0u, v);
    v += dpp_row<0x118>(0u, v);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

// Segmented inclusive OR across a wave with DPP moves only: lane l ORs the
// values of lanes [l - d, l] of its run (d = distance to the run's first
// lane; runs are contiguous).  Shifts 1, 2, 4, 8 within rows of 16 lanes,
// then row_bcast:15 (lane 15 / 47 into rows 1 / 3) and row_bcast:31 (lane 31
// into rows 2 and 3), each taken only when its source lane is in the run.
// All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_seg_or_dpp(uint32_t v, uint32_t d, uint32_t l) {
    uint32_t t;
    t = dpp_row<0x111>(0u, v);
    v |= d >= 1 ? t : 0u;
    t = dpp_row<0x112>(0u, v);
    v |= d >= 2 ? t : 0u;
    t = dpp_row<0x114>(0u, v);
    v |= d >= 4 ? t : 0u;
    t = dpp_row<0x118>(0u, v);
    v |= d >= 8 ? t : 0u;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v |= d >= (l & 15u) + 1u ? t : 0u;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    v |= (l >= 32u && d >= l - 31u) ? t : 0u;
    return v;
}

// ORs bit (1 << (x & 31)) into word x >> 5 of an LDS bitmap for one value
// per lane, all 64 lanes active, values ascending across the lanes: lanes
// whose values share a word OR their bits together first (a run of equal
// words is contiguous), and only the run's last lane issues the LDS atomic.
// A wave with no shared word issues one atomic per lane as usual.
__device__ __forceinline__ void wave_or_sorted_bits(uint32_t *bm, uint32_t x) {
    const uint32_t l = (uint32_t)lane();
    const uint32_t w = x >> 5;
    const uint32_t bit = 1u << (x & 31u);
    // the word of the lane below (wave_shr:1; lane 0 gets ~0, never a word)
    const uint32_t wprev = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)w, 0x138, 0xF, 0xF, false);
    const uint64_t cont = __ballot(l > 0 && wprev == w);
    if (cont == 0) {
        atomicOr(&bm[w], bit);
        return;
    }
    const uint64_t starts = ~cont;  // bit 0 is always set
    const uint64_t upto = starts & (l == 63u ? ~0ull : ((2ull << l) - 1ull));
    const uint32_t s = 63u - (uint32_t)__clzll((long long)upto);
    const uint32_t v = wave_seg_or_dpp(bit, l - s, l);
    const bool last = l == 63u || !((cont >> (l + 1u)) & 1ull);
    if (last) atomicOr(&bm[w], v);
}

// Number of set bits of m in the lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One step of ordered bitmap emission by a wave: lane l holds bitmap word
// `word` (64 consecutive words, the first one's columns start at cb).  The
// set bits are written as columns, in ascending order, to out[pos ...];
// positions at or beyond lim are dropped.  A step with at least `dense` set
// bits is written one lane per bit: two words per iteration, a ballot places
// each set lane, every store instruction is one contiguous run.  Sparser
// steps are written one lane per word, each lane walking its word's bits from
// its prefix position.  Returns the step's number of set bits.  All 64 lanes
// must be active.
__device__ __forceinline__ uint32_t emit_step(uint32_t word, uint32_t cb, uint64_t pos, uint32_t *out,
                                              uint64_t lim, uint32_t dense) {
    const uint32_t l = (uint32_t)lane();
    const uint32_t pc = (uint32_t)__popc(word);
    const uint32_t inc = wave_incl_sum_dpp(pc);
    const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    if (st == 0) return 0;
    if (pos + st > lim) {
        // only if the row's slot is too small (the passes disagreed): every
        // store is checked
        uint32_t w = word;
        uint64_t q = pos + (inc - pc);
        const uint32_t c0 = cb + 32u * l;
        while (w) {
            if (q < lim) out[q] = c0 + (uint32_t)__builtin_ctz(w);
            ++q;
            w &= w - 1u;
        }
        return st;
    }
    // the step fits: stores relative to one uniform (scalar) base, no checks
    const uint64_t up = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pos >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pos);
    uint32_t *o = out + up;
    if (st >= dense) {
        uint32_t p = 0;
        const uint32_t sh = l & 31u;
        for (uint32_t k = 0; k < 32; ++k) {
            const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)word, (int)(2 * k));
            const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)word, (int)(2 * k + 1));
            if ((w0 | w1) == 0) continue;
            const uint32_t v = l < 32 ? w0 : w1;
            const bool set = (v >> sh) & 1u;
            const uint64_t m = __ballot(set);
            if (set) *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(o) + 4u * (p + lanes_below(m))) =
                         cb + 64u * k + l;
            p += (uint32_t)__popcll(m);
        }
    } else {
        uint32_t w = word;
        uint32_t boff = (inc - pc) * 4u;
        const uint32_t c0 = cb + 32u * l;
        while (w) {
            *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(o) + boff) = c0 + (uint32_t)__builtin_ctz(w);
            boff += 4u;
            w &= w - 1u;
        }
    }
    return st;
}

// Staged emission: each lane writes its words' columns into the wave's
// staging area `sg` (STAGE_WORDS words), then the wave copies them out in
// order, 64 consecutive columns per store instruction.  Scattered per-lane
// stores would each touch their own cache line; the staged copy writes whole
// lines.  When the columns average more than four per lane, lanes write
// about 16 apart, which would put all 64 lanes on a few LDS banks, so column
// t then sits at sg[t + t / 32] (one padding word per 32); sparser rounds
// stage unpadded.
constexpr uint32_t STAGE_WORDS = 2048 + 64;
// Lane l copies columns l, l + 64, ... of n: column 64k + l sits at src[KS * k]
// (src = the lane's first slot; KS = 66 padded, 64 unpadded) and goes to byte
// 4 * (64k + l) of dst (a uniform pointer, so stores use a 32-bit offset).
template <uint32_t KS>
__device__ __forceinline__ void staged_copy_out(const uint32_t *src, char *dst, uint32_t l, uint32_t n) {
    const uint32_t full = n / WAVE;
    uint32_t k = 0, b = 4u * l;
    for (; k + 4 <= full; k += 4, b += 4u * 4u * WAVE) {
        const uint32_t v0 = src[KS * k], v1 = src[KS * (k + 1)], v2 = src[KS * (k + 2)], v3 = src[KS * (k + 3)];
        *reinterpret_cast<uint32_t *>(dst + b) = v0;
        *reinterpret_cast<uint32_t *>(dst + b + 4u * WAVE) = v1;
        *reinterpret_cast<uint32_t *>(dst + b + 8u * WAVE) = v2;
        *reinterpret_cast<uint32_t *>(dst + b + 12u * WAVE) = v3;
    }
    for (; k < full; ++k, b += 4u * WAVE) *reinterpret_cast<uint32_t *>(dst + b) = src[KS * k];
    if (WAVE * k + l < n) *reinterpret_cast<uint32_t *>(dst + b) = src[KS * k];
}

// Writes the set bits of 256 bitmap words as columns at out[pos ...] (never
// at or past lim): lane l holds four words x[0..3], bit b of x[e] being
// column cw[e] + b; the words are in column order by (lane, e).  One wave
// scan per round; a round whose columns fit the staging area goes through
// it, a denser one is written two words at a time with a lane per bit (64
// consecutive columns per ballot, stores in order).  Returns the round's set
// bits.
__device__ __forceinline__ uint32_t emit_round4_staged(const uint32_t (&x)[4], const uint32_t (&cw)[4], uint64_t pos,
                                                       uint32_t *out, uint64_t lim, uint32_t *sg) {
    const uint32_t l = (uint32_t)lane();
    const uint32_t pc = (uint32_t)(__popc(x[0]) + __popc(x[1]) + __popc(x[2]) + __popc(x[3]));
    const uint32_t inc = wave_incl_sum_dpp(pc);
    const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    if (st == 0) return 0;
    const uint64_t up = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pos >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pos);
    char *dst = reinterpret_cast<char *>(out + up);
    const uint32_t room = (uint32_t)(lim > up ? lim - up : 0);
    const uint32_t n = st < room ? st : room;
    if (st <= STAGE_WORDS - WAVE) {
        const bool padded = st > 4u * WAVE;  // uniform
        uint32_t q = inc - pc;
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            uint32_t w = x[e];
            while (w) {
                sg[padded ? q + (q >> 5) : q] = cw[e] + (uint32_t)__builtin_ctz(w);
                ++q;
                w &= w - 1u;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (padded) staged_copy_out<66>(sg + l + (l >> 5), dst, l, n);
        else staged_copy_out<64>(sg + l, dst, l, n);
        __builtin_amdgcn_wave_barrier();
    } else {
        // words (j, 2h) and (j, 2h + 1): lanes 0..31 take the first word's
        // bits, lanes 32..63 the second's
        const uint32_t sh = l & 31u;
        uint32_t p = 0;
        for (uint32_t j = 0; j < WAVE; ++j) {
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)x[2 * h], (int)j);
                const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)x[2 * h + 1], (int)j);
                if ((w0 | w1) == 0) continue;
                const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)cw[2 * h], (int)j);
                const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)cw[2 * h + 1], (int)j);
                const uint32_t v = l < 32 ? w0 : w1;
                const bool set = (v >> sh) & 1u;
                const uint64_t m = __ballot(set);
                const uint32_t at = p + lanes_below(m);
                if (set && at < n) *reinterpret_cast<uint32_t *>(dst + 4u * at) = (l < 32 ? b0 : b1) + sh;
                p += (uint32_t)__popcll(m);
            }
        }
    }
    return st;
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
