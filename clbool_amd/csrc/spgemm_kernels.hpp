// spgemm_kernels.hpp -- device side of the boolean hash SpGEMM.
//
// Restates the behaviour of clBool's dcsr::matrix_multiplication_hash
// (src/dcsr/dcsr_matrix_multiplication_hash.cpp:51-119 with the kernels of
// src/cl/count_workload.cl and src/cl/hash/*.cl) for CDNA4:
//
//  * B rows are found through a dense row-range map (one 8-byte load per A
//    nonzero) instead of a binary search over b_rows (count_workload.cl:8-27).
//  * Rows are binned by their intermediate-product count P (and, in the
//    numeric pass, by their exact output size N) as in hash.cpp:27-36, but
//    the bins are: TINY (one wave per row, register bitonic sort), HASH with
//    LDS tables of 2^7..2^14 slots (compact, then LDS bitonic sort of the
//    N live entries only), and BITMAP (an LDS bitmap over a column window;
//    the set bits come out in column order, so nothing is sorted).  The
//    reference's single-workgroup global-memory table plus global bitonic
//    sort for long rows (hash_global.cl:104-188) has no counterpart.
//  * Products of a row are enumerated load-balanced: B rows of >= 256
//    entries by the whole workgroup, of >= 64 by one wave, shorter ones by
//    lane-level expansion over a wave prefix sum.
#pragma once

#include "clb_internal.hpp"
#include "device_utils.hpp"

namespace clb {
namespace spg {

// Bin ids.
constexpr u32 NBINS = 32;
constexpr u32 BIN_EMPTY = 0;
constexpr u32 BIN_TINY = 1;
constexpr u32 BIN_HASH0 = 2;     // 2 + (log2 slots - 7)
constexpr u32 HASH_MIN_BITS = 7;
constexpr u32 HASH_MAX_BITS = 14;
constexpr u32 BIN_ESC0 = 10;     // 10 + (ceil(log2 P) - ESC_MIN_BITS): LDS expand-sort-compress rows
constexpr u32 ESC_MIN_BITS = 11;
constexpr u32 ESC_MAX_BITS = 13;
constexpr u32 BIN_BITMAP0 = 13;  // 13 + (31 - floor(log2 P)), clamped to 31
constexpr u32 BIG_ROW = 256;     // B rows at least this long are shared by the workgroup
constexpr u32 LONG_ROW = 64;     // B rows at least this long are shared by a wave

struct Mats {
    const u32 *a_rpt;
    const u32 *a_rows;
    const u32 *a_cols;
    const uint2 *brange;  // B row id -> [begin, end) into b_cols
    const u32 *b_cols;
    u32 mapn;             // entries of brange
    u32 row0;             // first A stored row of the block
    u32 ncols;            // columns of B (and C)
    u32 emit;             // bitmap emission: lane-per-bit threshold (EMIT_DEFAULT when 0)
    u32 sparse_bits;      // bitmap emission: set bits per 1024 words below which lanes write their own words
    u32 combine;          // bitmap gather: OR bits of one word within a wave before the LDS atomic
    // Column slices of the bitmap kernels (nsl > 1 when ncols exceeds one
    // window of 2^split_bits words): splits[k * (nsl + 1) + s] is the
    // position in b_cols of B row k's first column in slice s (s = nsl: the
    // row's end), so a slice reads only its own part of each B row; all
    // entries of a row B does not store are zero.  A kernel with a window
    // 2^d times as wide takes every 2^d-th boundary.
    const u32 *splits;
    u32 nsl;
    u32 split_bits;
    // Hypersparse B (few stored rows): one bit per B row id, set for stored
    // rows.  A lookup tests the bit (a 2 MB table for 2^24 rows stays in L2)
    // and reads brange only for stored rows; brange is written only there.
    // nullptr: brange is complete.
    const u32 *bpres;
};

// [begin, end) of B row k in b_cols ({0, 0} for absent rows).
__device__ __forceinline__ uint2 b_range(const Mats &g, u32 k) {
    if (k >= g.mapn) return make_uint2(0, 0);
    if (g.bpres && !((g.bpres[k >> 5] >> (k & 31)) & 1u)) return make_uint2(0, 0);
    return g.brange[k];
}
constexpr u32 NOSLICE = 0xFFFFFFFFu;

// {first position, length} in b_cols of what the A nonzero with column k
// reads: the whole B row (slice == NOSLICE) or its part in fine slices
// [slice * sstep, (slice + 1) * sstep) through the split table, which is
// zero for rows B does not store, so that path needs no brange lookup.
// k == EMPTY gives {0, 0}.
__device__ __forceinline__ uint2 b_part(const Mats &g, u32 k, u32 slice, u32 sstep) {
    if (slice == NOSLICE) {
        const uint2 rr = b_range(g, k);
        return make_uint2(rr.x, rr.y - rr.x);
    }
    if (k >= g.mapn) return make_uint2(0, 0);
    const u32 *sp = g.splits + (u64)k * (g.nsl + 1);
    const u32 f0 = slice * sstep, f1 = f0 + sstep < g.nsl ? f0 + sstep : g.nsl;
    const u32 b0 = sp[f0];
    return make_uint2(b0, sp[f1] - b0);
}

// Lookahead over the slices of one row: k0 is this lane's A column in the
// row's first chunk (EMPTY past the row's end) and part its B part in the
// slice being gathered; the gather of a slice leaves the next slice's part
// in it.
struct RowAhead {
    u32 k0;
    uint2 part;
};

// Bitmap emission.  Each wave owns a contiguous range of the bitmap's words
// and walks it 64 words at a time (dev::emit_step): a step whose 64 words
// hold at least `emit` set bits is written one lane per bit with coalesced
// stores, sparser steps one lane per word.
constexpr u32 EMIT_DEFAULT = 1024;

struct BinCfg {
    u32 ncols;
    u32 tiny_max;
    u32 hash_max_bits;
    u32 bitmap_min;
    u32 bitmap_ratio;
    u32 bm_bits;   // log2 of the bitmap window in 32-bit words (numeric and one-pass kernels)
    u32 sym_bits;  // the same for the symbolic kernel
    // rows the hash rule sends to a bitmap that would need several column
    // windows go to the LDS expand-sort-compress kernel instead when they
    // have at most esc_max products (0: never)
    u32 esc_max;
    u32 multi_slice;  // ncols spans more than one bitmap window
    u32 tables_over_slices;  // multi_slice: the slot-ratio rule does not send rows to a bitmap
    // the ESC SpGEMM (dcsr::matrix_multiplication): tiny, LDS ESC up to
    // 2^ESC_MAX_BITS products, longer rows to the global sort (BIN_BITMAP0)
    u32 esc_only;
};

__device__ __forceinline__ u32 ceil_log2(u32 x) { return x <= 1 ? 0 : 32 - __clz(x - 1); }

// P: products of the row; key: the number of distinct columns the table
// must hold (min(P, ncols) before the symbolic pass, N after it).
__device__ __forceinline__ u32 esc_bin(u32 P) {
    const u32 b = ceil_log2(P);
    return BIN_ESC0 + (b < ESC_MIN_BITS ? 0u : b - ESC_MIN_BITS);
}
__device__ __forceinline__ u32 classify(u32 P, u32 key, const BinCfg &c) {
    if (P == 0 || key == 0) return BIN_EMPTY;
    if (P <= c.tiny_max) return BIN_TINY;
    if (c.esc_only) return P <= (1u << ESC_MAX_BITS) ? esc_bin(P) : BIN_BITMAP0;
    u32 bits = ceil_log2(2 * key);
    if (bits < HASH_MIN_BITS) bits = HASH_MIN_BITS;
    // the slot-ratio rule (a bitmap no larger than ratio x the table costs
    // no more) assumes one window; with tables_over_slices a row that would
    // need several column slices takes the table instead
    const bool ratio_rule = !(c.multi_slice && c.tables_over_slices);
    const bool bitmap = P >= c.bitmap_min || bits > c.hash_max_bits ||
                        (ratio_rule && (1ull << c.bm_bits) <= (u64)c.bitmap_ratio * (1ull << bits));
    if (!bitmap) return BIN_HASH0 + (bits - HASH_MIN_BITS);
    if (c.multi_slice && P <= c.esc_max) return esc_bin(P);
    const u32 lp = 31 - __clz(P);
    const u32 b = BIN_BITMAP0 + (31 - lp);
    return b < NBINS ? b : NBINS - 1;
}

// ---------------------------------------------------------------------------
// Product enumeration of one A row [js, je) by a workgroup of WG threads.
// `big` is LDS scratch of BIG_WORDS(WG) words, 8-byte aligned (unused when
// WG == 64).  f(col) is called once per product, possibly from divergent
// lanes; ff(col) instead of f for the products of full workgroup-path
// windows, where it is called by all 64 lanes of a wave at once with the
// columns ascending across the lanes.
//
// The A row is taken in chunks of WG nonzeros; each B row of the chunk goes
// one of three ways by its length:
//  * >= BIGMIN (workgroup path): the chunk's long rows are cut into windows
//    of UB * 64 columns and listed in LDS with the exclusive prefix of their
//    window counts; wave w takes windows w, w+NW, ... of the concatenation.
//    A full window is UB loads of 64 consecutive columns at one base address
//    (immediate offsets, no per-lane checks); only a row's last window needs
//    clamping.  The wave keeps a cursor into the list, with 64 list entries
//    cached one per lane (read with readlane), so moving to the next B row
//    needs no LDS access.  Windows are double-buffered: the loads of one are
//    in flight while the previous one's columns are applied (the kernels
//    with the largest LDS tables run one workgroup per CU).
//  * >= LONG_ROW (wave path): one wave per row, UNROLL loads per lane.
//  * shorter (lane path): the wave's short rows are concatenated and each
//    lane finds its row by a binary search over the wave prefix.
constexpr int UNROLL = 4;
__host__ __device__ constexpr int BIG_WORDS(int WG) { return 3 * WG + 2 * (WG / 64 + 2); }

template <int WG, int UB, u32 BIGMIN, bool AHEAD = false, typename F, typename FF>
__device__ __forceinline__ void for_each_product(const Mats &g, u32 js, u32 je, u32 *big, F &&f, FF &&ff,
                                                 u32 slice = NOSLICE, u32 sstep = 1, RowAhead *ah = nullptr,
                                                 u32 next_slice = NOSLICE) {
    const int l = dev::lane();
    constexpr int NW = WG / dev::WAVE;
    constexpr u32 WIN = UB * 64;
    // AHEAD (a metadata pipeline, for kernels whose occupancy LDS limits:
    // it costs 7 to 18 VGPRs): `part` holds the B part of this lane's nonzero in
    // the current chunk and `knext` the A column of its nonzero in the next
    // chunk.  Before its gather, a chunk issues the loads of the next chunk's
    // part and of the column two chunks ahead (the first chunk, given a
    // RowAhead, also the next slice's first part), so that their latency
    // overlaps the gather's loads instead of preceding them.
    auto col_at = [&](u64 j) { return j < je ? g.a_cols[j] : EMPTY; };
    uint2 part = make_uint2(0u, 0u);
    u32 knext = EMPTY;
    if constexpr (AHEAD) {
        part = ah ? ah->part : b_part(g, col_at((u64)js + threadIdx.x), slice, sstep);
        knext = col_at((u64)js + WG + threadIdx.x);
    }
    for (u32 base = js; base < je; base += WG) {
        u32 bx, len;
        if constexpr (AHEAD) {
            bx = part.x;
            len = part.y;
            if (ah && base == js && next_slice != NOSLICE) ah->part = b_part(g, ah->k0, next_slice, sstep);
            if ((u64)base + WG < je) {
                part = b_part(g, knext, slice, sstep);
                knext = col_at((u64)base + 2 * WG + threadIdx.x);
            }
        } else {
            const uint2 p = b_part(g, col_at((u64)base + threadIdx.x), slice, sstep);
            bx = p.x;
            len = p.y;
        }
        // list entry e: big[3e] = first B position, big[3e+1] = end,
        // big[3e+2] = windows before it.  Window counts of distinct B rows
        // sum to less than 2^32, so the packed scan never carries into the
        // count.
        u32 nb = 0, nwin = 0;
        if (NW > 1) {
            const bool isb = len >= BIGMIN;
            u64 *red64 = reinterpret_cast<u64 *>(big + 3 * WG);
            u64 all;
            const u64 ex =
                dev::block_excl_scan<u64, WG>(isb ? ((1ull << 32) | ((len + WIN - 1) / WIN)) : 0ull, red64, all);
            nb = (u32)(all >> 32);
            nwin = (u32)all;
            if (isb) {
                const u32 e = (u32)(ex >> 32);
                big[3 * e] = bx;
                big[3 * e + 1] = bx + len;
                big[3 * e + 2] = (u32)ex;
                len = 0;
            }
        }
        // wave-shared long rows
        unsigned long long lm = __ballot(len >= LONG_ROW);
        while (lm) {
            const int s = __ffsll(lm) - 1;
            lm &= lm - 1;
            const u32 sb = __shfl(bx, s, dev::WAVE);
            const u32 sl = __shfl(len, s, dev::WAVE);
            for (u32 q = l; q < sl; q += UNROLL * dev::WAVE) {
                u32 c[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    const u32 qq = q + u * dev::WAVE;
                    c[u] = qq < sl ? g.b_cols[sb + qq] : EMPTY;
                }
#pragma unroll
                for (int u = 0; u < UNROLL; ++u)
                    if (c[u] != EMPTY) f(c[u]);
            }
        }
        // lane expansion of short rows
        const u32 sl = len < LONG_ROW ? len : 0;
        const u32 inc = dev::wave_incl_scan(sl);
        const u32 wtot = __shfl(inc, dev::WAVE - 1, dev::WAVE);
        const u32 exc = inc - sl;
        for (u32 p0 = 0; p0 < wtot; p0 += UNROLL * dev::WAVE) {
            u32 c[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const u32 p = p0 + u * dev::WAVE + l;
                int s = 0;
                u32 es = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const u32 e = __shfl(exc, s + step, dev::WAVE);
                    if (e <= p) {
                        s += step;
                        es = e;
                    }
                }
                const u32 sb = __shfl(bx, s, dev::WAVE);
                c[u] = p < wtot ? g.b_cols[sb + (p - es)] : EMPTY;
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
                if (c[u] != EMPTY) f(c[u]);
        }
        if (NW > 1 && nb) {
            __syncthreads();  // the list is complete
            const u32 w = __builtin_amdgcn_readfirstlane(threadIdx.x / dev::WAVE);
            // list entries [cbase, cbase + 64) cached one per lane
            u32 cbase = 0, cbx = 0, cend = 0, cpre = 0;
            auto load_cache = [&](u32 b0) {
                cbase = b0;
                const u32 q = b0 + (u32)l;
                cbx = q < nb ? big[3 * q] : 0u;
                cend = q < nb ? big[3 * q + 1] : 0u;
                cpre = q < nb ? big[3 * q + 2] : nwin;
            };
            auto lane_val = [](u32 v, u32 lane) { return (u32)__builtin_amdgcn_readlane((int)v, (int)lane); };
            load_cache(0);
            // wave-uniform cursor: entry e holds windows [pre, next)
            u32 e = 0, pre = 0;
            u32 eb = lane_val(cbx, 0), ee = lane_val(cend, 0), next = lane_val(cpre, 1);
            // loads window k into c; returns its number of columns (0: none).
            // Every call issues exactly UB loads (past a row's end, or past
            // the last window, lanes re-read a column of the current entry,
            // and apply masks them off): with a fixed number of loads in
            // flight the compiler waits for one buffer only while the other
            // buffer's loads are still outstanding, instead of draining all
            // loads before each apply.
            auto fetch = [&](u32 k, u32(&c)[UB]) -> u32 {
                u32 start = eb, n = 0;
                if (k < nwin) {
                    while (k >= next) {
                        ++e;
                        if (e + 1 - cbase >= (u32)dev::WAVE) load_cache(e);
                        const u32 s = e - cbase;
                        pre = next;
                        eb = lane_val(cbx, s);
                        ee = lane_val(cend, s);
                        next = lane_val(cpre, s + 1);
                    }
                    start = eb + (k - pre) * WIN;
                    n = ee - start < WIN ? ee - start : WIN;
                }
                const u32 *src = g.b_cols + start;
                const u32 last = n ? n - 1 : 0u;  // entries hold at least one column
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const u32 o = (u32)(u * 64 + l);
                    c[u] = src[o < last ? o : last];
                }
                return n;
            };
            auto apply = [&](const u32(&c)[UB], u32 n) {
                if (n == WIN) {
                    // all 64 lanes hold consecutive columns of one B row
#pragma unroll
                    for (int u = 0; u < UB; ++u) ff(c[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < UB; ++u)
                        if ((u32)(u * 64 + l) < n) f(c[u]);
                }
            };
            u32 ca[UB], cb[UB];
            u32 na = fetch(w, ca), nb2 = 0;
            for (u32 k = w; k < nwin; k += 2 * NW) {
                nb2 = fetch(k + NW, cb);
                apply(ca, na);
                if (k + NW >= nwin) break;
                na = fetch(k + 2 * NW, ca);
                apply(cb, nb2);
            }
            __syncthreads();  // before the next chunk rewrites the list
        }
    }
}

// The same with one callable for every product.
template <int WG, int UB, u32 BIGMIN, bool AHEAD = false, typename F>
__device__ __forceinline__ void for_each_product(const Mats &g, u32 js, u32 je, u32 *big, F &&f,
                                                 u32 slice = NOSLICE, u32 sstep = 1, RowAhead *ah = nullptr,
                                                 u32 next_slice = NOSLICE) {
    for_each_product<WG, UB, BIGMIN, AHEAD>(g, js, je, big, f, f, slice, sstep, ah, next_slice);
}

// ---------------------------------------------------------------------------
// Per-row statistics of rows [row0, row0 + nrows) of A over its B rows,
// load-balanced by nonzeros: workgroup b takes A nonzeros [j0, j0 + 256) of
// the rows' range, finds their rows from a copy of a_rpt in LDS, reduces per
// row in LDS and merges into the per-row arrays with one global atomic per
// row and workgroup (a row spread over several workgroups is merged from
// each).  A hub row of A with many nonzeros thus costs no more than as many
// short rows, where one thread group per row would walk it serially.
//   prod[r] += products (must start at 0); total += all products;
//   with FEAT also mn/mx (min/max B-row length, start at ~0 / 0) and fc/lc
//   (first column of any of its B rows / last column, start at ~0 / 0).
constexpr u32 RS_WG = 256;
template <bool FEAT>
__global__ __launch_bounds__(RS_WG) void k_row_stats(Mats g, u32 nrows, u32 *__restrict__ prod,
                                                     u64 *__restrict__ total, u32 *__restrict__ mn,
                                                     u32 *__restrict__ mx, u32 *__restrict__ fc,
                                                     u32 *__restrict__ lc) {
    __shared__ u32 rs[RS_WG + 2];
    __shared__ u32 sp[RS_WG + 1], smn[RS_WG + 1], smx[RS_WG + 1], sfc[RS_WG + 1], slc[RS_WG + 1];
    __shared__ u64 red[RS_WG / 64 + 1];
    const u32 rbeg = g.row0, rend = g.row0 + nrows;
    const u32 jbase = g.a_rpt[rbeg], jend = g.a_rpt[rend];
    const u64 j0 = (u64)jbase + (u64)blockIdx.x * RS_WG;
    if (j0 >= jend) return;  // uniform: the grid covers all of A's nonzeros
    // last row r in [rbeg, rend) with a_rpt[r] <= j0: rounds of a 256-way
    // split of [lo, hi), one probe per thread (a_rpt[lo] <= j0 < a_rpt[hi]
    // throughout); three rounds cover 2^24 rows, where one thread's binary
    // search would be a chain of 24 dependent loads
    u32 lo = rbeg, hi = rend;
    while (hi - lo > 1) {
        const u32 step = (hi - lo + RS_WG - 1) / RS_WG;
        const u64 c = (u64)lo + (u64)threadIdx.x * step;
        const bool ok = c < hi && g.a_rpt[c] <= j0;  // true for a prefix of the threads
        const u32 cnt = (u32)__syncthreads_count(ok);  // >= 1: the probe at lo holds
        lo += (cnt - 1) * step;
        hi = hi - lo < step ? hi : lo + step;
    }
    const u32 r0 = lo;
    // rows r0 .. r0 + nr - 1 can hold this workgroup's nonzeros (every stored
    // row has one, so at most RS_WG + 1 of them)
    const u32 nr = rend - r0 < RS_WG + 1 ? rend - r0 : RS_WG + 1;
    for (u32 t = threadIdx.x; t <= nr; t += RS_WG) rs[t] = g.a_rpt[r0 + t];
    for (u32 t = threadIdx.x; t < nr; t += RS_WG) {
        sp[t] = 0;
        if (FEAT) {
            smn[t] = 0xFFFFFFFFu;
            smx[t] = 0;
            sfc[t] = 0xFFFFFFFFu;
            slc[t] = 0;
        }
    }
    __syncthreads();
    const u64 j = j0 + threadIdx.x;
    u32 mine = 0;
    if (j < jend) {
        // local row: last t with rs[t] <= j
        u32 lo = 0, hi = nr;
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) / 2;
            if (rs[mid] <= j) lo = mid;
            else hi = mid;
        }
        const u32 k = g.a_cols[j];
        {
            const uint2 b = b_range(g, k);
            const u32 len = b.y - b.x;
            if (len) {
                mine = len;
                atomicAdd(&sp[lo], len);
                if (FEAT) {
                    atomicMin(&smn[lo], len);
                    atomicMax(&smx[lo], len);
                    atomicMin(&sfc[lo], g.b_cols[b.x]);
                    atomicMax(&slc[lo], g.b_cols[b.y - 1]);
                }
            }
        }
    }
    const u64 bs = dev::block_sum<u64, RS_WG>((u64)mine, red);
    if (threadIdx.x == 0 && bs) atomicAdd(reinterpret_cast<unsigned long long *>(total), (unsigned long long)bs);
    for (u32 t = threadIdx.x; t < nr; t += RS_WG) {
        // only rows that overlap this workgroup's nonzeros
        if (rs[t + 1] <= j0 || rs[t] >= j0 + RS_WG || rs[t] >= jend) continue;
        const u32 row = r0 + t - rbeg;
        if (sp[t]) atomicAdd(&prod[row], sp[t]);
        if (FEAT && sp[t]) {
            atomicMin(&mn[row], smn[t]);
            atomicMax(&mx[row], smx[t]);
            atomicMin(&fc[row], sfc[t]);
            atomicMax(&lc[row], slc[t]);
        }
    }
}

// Insert into an LDS open-addressing table of 2^BITS slots (linear probing,
// EMPTY marks a free slot).  Returns 1 if col was new, 0 if present, and 2
// if the table is full (only possible when a capacity guess was too small)
// or after `max_probe` probes.
template <int BITS>
__device__ __forceinline__ u32 hash_insert(u32 *table, u32 col, u32 max_probe) {
    constexpr u32 T = 1u << BITS;
    u32 h = dev::hash_slot(col, BITS);
    const u32 np = max_probe < T ? max_probe : T;
    for (u32 probe = 0; probe < np; ++probe) {
        // one LDS round trip per probe: the compare-and-swap returns what
        // the slot held (EMPTY: inserted; col: already there)
        const u32 cur = atomicCAS(&table[h], EMPTY, col);
        if (cur == EMPTY) return 1;
        if (cur == col) return 0;
        h = (h + 1) & (T - 1);
    }
    return 2;
}

// ---------------------------------------------------------------------------
// Pass modes shared by the row kernels.
//  SYM: count the distinct columns of each row into nnz_row (symbolic pass).
//  NUM: write each row's sorted columns at off[r]; nnz_row holds the exact
//       count from the symbolic pass (numeric pass).
//  ONE: single pass sized by a capacity cap[r]: write up to cap[r] sorted
//       columns at off[r] and the count into nnz_row[r]; a row with more
//       distinct columns than its capacity writes nothing, stores FAILED and
//       bumps *overflow (its exact re-run happens in a second launch).
constexpr int SYM = 0, NUM = 1, ONE = 2;
constexpr u32 FAILED = 0xFFFFFFFFu;

// Dense bitmap rows (windows of 2^14 or 2^15 words, every slice of the row)
// leave the gather kernel in global memory, each slice as its bitmap or, when
// at most a third of its words are nonzero, as (word index, word) pairs of
// those words, plus popcounts per 1024-word chunk; k_bitmap_emit expands them
// later into their final positions with coalesced stores.  store == nullptr
// disables this path.
constexpr u32 EMIT_CHUNK = 1024;  // words per emission chunk
// Slots are handed out in the order rows decide to dump (*count counts the
// claims); a row that finds all `capacity` slots taken is not dumped.
struct Dump {
    uint4 *store;      // [slot][nslices][W/4]
    u32 *chunk_cnt;    // [slot][nslices][W / EMIT_CHUNK] popcounts (zero for unused slots)
    u32 *list;         // [slot]: the dumped row's index in the bitmap range of perm
    u32 *count;        // slots claimed (may exceed capacity)
    u32 *isdense;      // [row] 1 if the row was dumped
    u32 dense_min;     // dump rows with at least this many columns in their first window
    u32 capacity;      // slots in store
    u32 *fmt;          // [slot][nslices]: 0 = the slice's bitmap, else 1 + the number of
                       // (word index, word) pairs of its nonzero words
};

struct Out {
    u32 *nnz_row;
    const u64 *off;
    u32 *cols;
    const u32 *cap;
    u32 *overflow;  // SYM/NUM: bit flags of inconsistencies; ONE: count of failed rows
    Dump dump;
    // ONE: a hash insertion that probes this many slots fails the row (0:
    // no limit).  With a table of at least twice the capacity, a long probe
    // sequence means the row holds more distinct columns than its capacity
    // (the prediction was too small); stopping early keeps a badly sized row
    // from probing a full table once per product.  The row is re-run exactly.
    u32 probe_limit;
};
constexpr u32 ONE_PASS_PROBES = 256;

template <int MODE>
__device__ __forceinline__ u32 row_limit(const Out &o, u32 r) {
    return MODE == NUM ? o.nnz_row[r] : (MODE == ONE ? o.cap[r] : 0u);
}

// TINY rows (P <= 64): one wave per row, one product per lane, register
// bitonic sort across the wave, duplicates removed by a neighbour compare.
template <int WAVES, int MODE>
__global__ __launch_bounds__(WAVES * 64) void k_tiny(Mats g, const u32 *__restrict__ perm, u32 count, Out o) {
    const u32 idx = blockIdx.x * WAVES + dev::wave_id();
    if (idx >= count) return;
    const int l = dev::lane();
    const u32 r = perm[idx];
    const u32 i = g.row0 + r;
    const u32 js = g.a_rpt[i], je = g.a_rpt[i + 1];
    u32 val = EMPTY;
    u32 run = 0;
    for (u32 base = js; base < je && run < 64; base += 64) {
        const u32 j = base + l;
        u32 bx = 0, len = 0;
        if (j < je) {
            const uint2 rr = b_range(g, g.a_cols[j]);
            bx = rr.x;
            len = rr.y - rr.x;
        }
        const u32 inc = dev::wave_incl_scan(len);
        const u32 tot = __shfl(inc, 63, 64);
        const u32 exc = inc - len;
        const u32 p = (u32)l - run;  // wraps when l < run
        int s = 0;
        u32 es = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const u32 e = __shfl(exc, s + step, 64);
            if (e <= p) {
                s += step;
                es = e;
            }
        }
        const u32 sb = __shfl(bx, s, 64);
        if ((u32)l >= run && p < tot) val = g.b_cols[sb + (p - es)];
        run += tot;
    }
    // bitonic sort of 64 values across the wave
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const u32 x = __shfl_xor(val, j, 64);
            const bool up = (l & k) == 0;
            const bool lower = (l & j) == 0;
            const u32 mn = x < val ? x : val, mx = x < val ? val : x;
            val = (lower == up) ? mn : mx;
        }
    }
    const u32 prev = __shfl_up(val, 1, 64);
    const bool keep = val != EMPTY && (l == 0 || prev != val);
    const unsigned long long km = __ballot(keep);
    const u32 n = (u32)__popcll(km);
    if (MODE == SYM) {
        if (l == 0) o.nnz_row[r] = n;
        return;
    }
    const u32 lim = row_limit<MODE>(o, r);
    if (MODE == ONE) {
        if (l == 0) o.nnz_row[r] = n <= lim ? n : FAILED;
        if (n > lim) {
            if (l == 0) atomicAdd(o.overflow, 1u);
            return;
        }
    }
    const u32 pos = (u32)__popcll(km & ((1ull << l) - 1ull));
    if (keep && pos < lim) o.cols[o.off[r] + pos] = val;
}

// ---------------------------------------------------------------------------
// HASH rows: LDS table of 2^BITS slots per workgroup.
template <int BITS, int WG, int MODE>
__global__ __launch_bounds__(WG) void k_hash(Mats g, const u32 *__restrict__ perm, Out o) {
    constexpr u32 T = 1u << BITS;
    __shared__ __align__(16) u32 table[T];
    __shared__ __align__(8) u32 big[BIG_WORDS(WG)];
    __shared__ u32 red[WG / 64 + 1];
    const u32 r = perm[blockIdx.x];
    const u32 i = g.row0 + r;
    const u32 js = g.a_rpt[i], je = g.a_rpt[i + 1];
    if (T >= 4 * WG) {
        uint4 *t4 = reinterpret_cast<uint4 *>(table);
        for (u32 t = threadIdx.x; t < T / 4; t += WG) t4[t] = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
    } else {
        for (u32 t = threadIdx.x; t < T; t += WG) table[t] = EMPTY;
    }
    __syncthreads();
    u32 mine = 0, full = 0;
    const u32 max_probe = MODE == ONE && o.probe_limit ? o.probe_limit : T;
    auto ins = [&](u32 col) {
        const u32 rc = hash_insert<BITS>(table, col, max_probe);
        mine += rc == 1;
        full |= rc == 2;
    };
    for_each_product<WG, 8, BIG_ROW>(g, js, je, big, ins);
    __syncthreads();
    if (MODE == SYM) {
        if (full) atomicOr(o.overflow, 1u);
        const u32 n = dev::block_sum<u32, WG>(mine, red);
        if (threadIdx.x == 0) o.nnz_row[r] = n;
        return;
    }
    const u32 lim = row_limit<MODE>(o, r);
    if (MODE == ONE) {
        const u32 n = dev::block_sum<u32, WG>(mine, red);
        const u32 anyfull = dev::block_sum<u32, WG>(full, red);
        if (n > lim || anyfull) {
            if (threadIdx.x == 0) {
                o.nnz_row[r] = FAILED;
                atomicAdd(o.overflow, 1u);
            }
            return;
        }
        if (threadIdx.x == 0) o.nnz_row[r] = n;
    } else if (full) {
        atomicOr(o.overflow, 1u);
    }
    // compact the live slots to the front of the table
    // (strided slots: consecutive lanes read consecutive words, no bank
    // conflicts; the packed order does not matter, the sort follows)
    constexpr u32 PER = T / WG > 0 ? T / WG : 1;
    u32 v[PER];
    u32 c = 0;
#pragma unroll
    for (u32 q = 0; q < PER; ++q) {
        const u32 t = q * WG + threadIdx.x;
        v[q] = t < T ? table[t] : EMPTY;
        c += v[q] != EMPTY;
    }
    u32 n;
    u32 pos = dev::block_excl_scan<u32, WG>(c, red, n);
#pragma unroll
    for (u32 q = 0; q < PER; ++q)
        if (v[q] != EMPTY) table[pos++] = v[q];
    const u32 *sorted = table;
    if (((n + 63u) & ~63u) <= T / 2) {
        // the upper half of the table is the merge scratch
        __syncthreads();
        sorted = dev::lds_merge_sort<WG>(table, table + T / 2, n);
    } else {
        // only when the table overflowed (the passes disagreed)
        const u32 n2 = n <= 1 ? 1 : (1u << (32 - __clz(n - 1)));
        for (u32 t = n + threadIdx.x; t < n2; t += WG) table[t] = EMPTY;
        __syncthreads();
        dev::lds_bitonic_sort<WG>(table, n2);
    }
    // never write past the row's slot, even if the passes disagreed
    const u32 w = n < lim ? n : lim;
    if (MODE == NUM && n != lim && threadIdx.x == 0) atomicOr(o.overflow, 2u);
    u32 *out = o.cols + o.off[r];
    for (u32 t = threadIdx.x; t < w; t += WG) out[t] = sorted[t];
}

// ---------------------------------------------------------------------------
// ESC rows (expand, sort, compress; the per-row scheme of bitonic_esc.cl:213-304):
// the row's P <= 2^CAPB products are appended to an LDS array (one LDS
// atomic per wave and call: the active lanes take consecutive slots by a
// ballot), sorted by dev::lds_merge_sort (wave bitonic sorts of 64-blocks,
// then merge-path rounds; the reference bitonic-sorts the whole array) and
// compressed: each thread keeps the values of its contiguous stretch that
// differ from their predecessor and writes them at its scanned position.
template <int CAPB, int WG, int MODE>
__global__ __launch_bounds__(WG) void k_esc(Mats g, const u32 *__restrict__ perm, Out o) {
    constexpr u32 CAP = 1u << CAPB;
    __shared__ __align__(16) u32 keys[CAP];
    __shared__ __align__(16) u32 scratch[CAP];
    __shared__ __align__(8) u32 big[BIG_WORDS(WG)];
    __shared__ u32 red[WG / 64 + 1];
    __shared__ u32 cnt;
    const u32 r = perm[blockIdx.x];
    const u32 i = g.row0 + r;
    const u32 js = g.a_rpt[i], je = g.a_rpt[i + 1];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    auto append = [&](u32 col) {
        const unsigned long long m = __ballot(1);
        const u32 leader = (u32)__ffsll(m) - 1u;
        u32 base = 0;
        if ((u32)dev::lane() == leader) base = atomicAdd(&cnt, (u32)__popcll(m));
        base = __shfl(base, (int)leader, dev::WAVE);
        const u32 slot = base + dev::lanes_below(m);
        if (slot < CAP) keys[slot] = col;
    };
    for_each_product<WG, 8, BIG_ROW>(g, js, je, big, append);
    __syncthreads();
    const u32 n = cnt < CAP ? cnt : CAP;  // the bin holds rows of at most CAP products
    const u32 *sorted = dev::lds_merge_sort<WG>(keys, scratch, n);
    const u32 per = (n + WG - 1) / WG;
    const u32 t0 = threadIdx.x * per < n ? threadIdx.x * per : n;
    const u32 t1 = t0 + per < n ? t0 + per : n;
    u32 c = 0;
    for (u32 t = t0; t < t1; ++t) c += (t == 0 || sorted[t] != sorted[t - 1]);
    u32 total;
    u32 pos = dev::block_excl_scan<u32, WG>(c, red, total);
    if (MODE == SYM) {
        if (threadIdx.x == 0) o.nnz_row[r] = total;
        return;
    }
    const u32 lim = row_limit<MODE>(o, r);
    if (MODE == ONE) {
        if (total > lim) {
            if (threadIdx.x == 0) {
                o.nnz_row[r] = FAILED;
                atomicAdd(o.overflow, 1u);
            }
            return;
        }
        if (threadIdx.x == 0) o.nnz_row[r] = total;
    } else if (total != lim && threadIdx.x == 0) {
        atomicOr(o.overflow, 2u);
    }
    u32 *out = o.cols + o.off[r];
    for (u32 t = t0; t < t1; ++t)
        if ((t == 0 || sorted[t] != sorted[t - 1]) && pos < lim) out[pos++] = sorted[t];
}

// ---------------------------------------------------------------------------
// BITMAP rows: an LDS bitmap of 2^WBITS words covers 32 * 2^WBITS columns;
// wider column ranges are processed in slices.
// One row (position idx of the bitmap bins) of the bitmap kernel.
// COMB: full windows OR their words within the wave first (Mats::combine;
// a separate instantiation, so the plain gather keeps its register budget).
template <int WBITS, int WG, int MODE, bool COMB>
__device__ __forceinline__ void bitmap_row(const Mats &g, const u32 *__restrict__ perm, u32 idx, const Out &o,
                                           u32 *bm, u32 *big, u32 *red) {
    constexpr u32 W = 1u << WBITS;
    constexpr u64 SPAN = 32ull * W;
    const u32 r = perm[idx];
    const u32 i = g.row0 + r;
    const u32 js = g.a_rpt[i], je = g.a_rpt[i + 1];
    const u32 nslices = (u32)(((u64)g.ncols + SPAN - 1) / SPAN);
    // the split table's slices are this window or a power-of-two fraction of it
    const bool use_splits = g.splits && g.nsl > 1 && (u32)WBITS >= g.split_bits;
    const u32 sstep = use_splits ? 1u << (WBITS - g.split_bits) : 1u;
    u32 mine = 0;
    u64 written = 0;
    bool dumping = false;  // uniform: the row goes to the dense-row emission
    u32 slot = 0;          // its slot in the dump
    u32 *out = MODE == SYM ? nullptr : o.cols + o.off[r];
    const u64 lim = row_limit<MODE>(o, r);
    // the metadata lookahead costs ~16 VGPRs: only where LDS already holds
    // occupancy to two 512-thread (or one 1024-thread) workgroups per CU
    constexpr bool AH = WBITS >= 14;
    RowAhead ah{EMPTY, make_uint2(0u, 0u)};
    if (AH && use_splits && nslices > 1) {
        if (threadIdx.x < je - js) ah.k0 = g.a_cols[js + threadIdx.x];
        ah.part = b_part(g, ah.k0, 0, sstep);
    }
    for (u32 s = 0; s < nslices; ++s) {
        const u32 sn = s + 1 < nslices ? s + 1 : NOSLICE;
        const u64 base = (u64)s * SPAN;
        {
            uint4 *bm4 = reinterpret_cast<uint4 *>(bm);
            for (u32 t = threadIdx.x; t < W / 4; t += WG) bm4[t] = make_uint4(0, 0, 0, 0);
        }
        __syncthreads();
        auto set1 = [&](u32 col) { atomicOr(&bm[col >> 5], 1u << (col & 31)); };
        auto setw = [&](u32 col) { dev::wave_or_sorted_bits(bm, col); };
        auto setd = [&](u32 col) {
            const u64 d = (u64)col - base;
            if (d < SPAN) atomicOr(&bm[(u32)(d >> 5)], 1u << (col & 31));
        };
        // a slice read through the split table holds only its own columns:
        // no range check (base < 2^32 and a multiple of 32)
        auto setu = [&](u32 col) { atomicOr(&bm[(col - (u32)base) >> 5], 1u << (col & 31)); };
        auto setdw = [&](u32 col) { dev::wave_or_sorted_bits(bm, (u32)((u64)col - base)); };
        constexpr int UBW = MODE == SYM ? 16 : 8;
        if (nslices == 1) {
            if constexpr (COMB) for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, set1, setw);
            else for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, set1);
        } else {
            // only the slice's part of each B row is read (with the split table)
            if constexpr (COMB) {
                if (use_splits)
                    for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, setu, setdw, s, sstep, &ah, sn);
                else for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, setd, NOSLICE, sstep);
            } else {
                if (use_splits) for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, setu, s, sstep, &ah, sn);
                else for_each_product<WG, UBW, LONG_ROW, AH>(g, js, je, big, setd, NOSLICE, sstep);
            }
        }
        __syncthreads();
        if (MODE == SYM) {
            const uint4 *bm4 = reinterpret_cast<const uint4 *>(bm);
            for (u32 t = threadIdx.x; t < W / 4; t += WG) {
                const uint4 v = bm4[t];
                mine += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
            }
        } else {
            // Each wave owns words [wv * WPW, (wv + 1) * WPW) of the window: in
            // round m, lane l takes the E consecutive words starting at
            // (m * 64 + l) * E (16-byte reads by consecutive lanes, no bank
            // conflicts).  One read of every round gives the counts (one
            // exchange of the wave totals gives the slice total for the dump
            // decision and each wave's offset); the dump re-reads every
            // round, the emission only the rounds that hold bits, with one
            // wave scan per round and each lane writing its own words'
            // columns.
            constexpr u32 NW = WG / dev::WAVE, WPW = W / NW, LPW = WPW / dev::WAVE;
            constexpr u32 E = LPW >= 4 ? 4 : LPW, R = LPW / E;
            const int l = dev::lane();
            const u32 wv = dev::wave_id();
            u32 *wb = bm + wv * WPW;
            auto words = [&](u32 m, u32(&x)[E]) {
                const u32 i0 = (m * dev::WAVE + (u32)l) * E;
                if constexpr (E == 4) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(wb + i0);
                    x[0] = v.x;
                    x[1] = v.y;
                    x[2] = v.z;
                    x[3] = v.w;
                } else if constexpr (E == 2) {
                    const uint2 v = *reinterpret_cast<const uint2 *>(wb + i0);
                    x[0] = v.x;
                    x[1] = v.y;
                } else {
                    x[0] = wb[i0];
                }
            };
            u32 c[R];
            u32 cnt = 0, nzw = 0;  // set bits and nonzero words of this lane
#pragma unroll
            for (u32 m = 0; m < R; ++m) {
                u32 x[E];
                words(m, x);
                c[m] = 0;
#pragma unroll
                for (u32 e = 0; e < E; ++e) {
                    c[m] += __popc(x[e]);
                    nzw += x[e] != 0u;
                }
                cnt += c[m];
            }
            // one exchange carries both wave totals: set bits (at most
            // 32 * WPW <= 2^16) above 12 bits of nonzero words (at most WPW
            // <= 2^11)
            static_assert(WPW <= 2048, "packed wave totals");
            const u32 wcnt = dev::wave_sum((cnt << 12) | nzw);
            if (l == 0) red[wv] = wcnt;
            __syncthreads();
            u32 woff = 0, tot = 0, nzoff = 0, nztot = 0;
            for (u32 v = 0; v < NW; ++v) {
                const u32 t = red[v] >> 12, z = red[v] & 0xFFFu;
                woff += v < wv ? t : 0u;
                tot += t;
                nzoff += v < wv ? z : 0u;
                nztot += z;
            }
            __syncthreads();
            if constexpr (WBITS >= 14) if (o.dump.store) {
                // Slice 0 decides for the whole row: a dense row has every
                // slice dumped (slice s at word s * W of its slot) and is
                // written later by the emission kernel.  A one-pass row over
                // its capacity already in its first window is dumped too: a
                // dumped row needs no temporary slot.
                if (s == 0) {
                    dumping = tot >= o.dump.dense_min || (MODE == ONE && tot > lim);  // tot is uniform
                    if (dumping) {
                        // claim a slot (big is free between gathers)
                        if (threadIdx.x == 0) big[0] = atomicAdd(o.dump.count, 1u);
                        __syncthreads();
                        slot = big[0];
                        __syncthreads();
                        dumping = slot < o.dump.capacity;
                    }
                }
                if (dumping) {
                    // a wave's range holds CPW dump chunks of RPC rounds each
                    constexpr u32 NCH = W / EMIT_CHUNK, CPW = WPW / EMIT_CHUNK, RPC = R / CPW;
                    static_assert(E == 4 && CPW * EMIT_CHUNK == WPW && RPC * CPW == R, "dump chunk layout");
                    const u64 sl = (u64)slot * nslices + s;
#pragma unroll
                    for (u32 k = 0; k < CPW; ++k) {
                        u32 cc = 0;
#pragma unroll
                        for (u32 m = k * RPC; m < (k + 1) * RPC; ++m) cc += c[m];
                        cc = dev::wave_sum(cc);
                        if (l == 0) o.dump.chunk_cnt[sl * NCH + wv * CPW + k] = cc;
                    }
                    // a slice with at most W / 3 nonzero words is dumped as
                    // (word index, word) pairs of those words (at most two
                    // thirds of the bitmap's bytes; W / 4 and W / 8 measured
                    // slower): one wave scan per round, at most four 8-byte
                    // stores per lane
                    const bool as_pairs = nztot <= W / 3;  // uniform
                    if (as_pairs) {
                        uint2 *pairs = reinterpret_cast<uint2 *>(o.dump.store + sl * (W / 4));
                        u32 q0 = nzoff;
#pragma unroll 1
                        for (u32 m = 0; m < R; ++m) {
                            u32 x[E];
                            u32 nz = 0;
                            if (c[m]) {
                                words(m, x);
#pragma unroll
                                for (u32 e = 0; e < E; ++e) nz += x[e] != 0u;
                            }
                            const u32 inc = dev::wave_incl_sum_dpp(nz);
                            const u32 rt = (u32)__builtin_amdgcn_readlane((int)inc, 63);
                            if (rt == 0) continue;  // uniform
                            if (nz) {
                                u32 q = q0 + (inc - nz);
                                const u32 i0 = wv * WPW + (m * dev::WAVE + (u32)l) * E;
#pragma unroll
                                for (u32 e = 0; e < E; ++e)
                                    if (x[e]) pairs[q++] = make_uint2(i0 + e, x[e]);
                            }
                            q0 += rt;
                        }
                    } else {
                        uint4 *dst = o.dump.store + sl * (W / 4) + wv * (WPW / 4);
                        const uint4 *src4 = reinterpret_cast<const uint4 *>(wb);
#pragma unroll 4
                        for (u32 m = 0; m < R; ++m) dst[m * dev::WAVE + (u32)l] = src4[m * dev::WAVE + (u32)l];
                    }
                    if (threadIdx.x == 0) o.dump.fmt[sl] = as_pairs ? nztot + 1u : 0u;
                    written += tot;
                    __syncthreads();  // before the next slice clears the bitmap
                    continue;
                }
            }
            if (MODE == ONE && written + tot > lim) {
                // the capacity was too small: nothing of the row is kept
                if (threadIdx.x == 0) {
                    o.nnz_row[r] = FAILED;
                    atomicAdd(o.overflow, 1u);
                }
                return;
            }
            if ((u64)tot * 1024u >= (u64)g.sparse_bits * W) {
                // a denser slice: 64 words per step (their stores land close
                // together), a lane per bit where a step is dense enough
                const u64 pos = written + woff;
                const u32 dense = g.emit ? g.emit : EMIT_DEFAULT;
                u64 done = 0;
                for (u32 q0 = 0; q0 < WPW; q0 += dev::WAVE) {
                    const u32 word = wb[q0 + l];
                    done += dev::emit_step(word, (u32)(base + 32ull * (wv * WPW + q0)), pos + done, out, lim, dense);
                }
            } else {
                u64 pos = written + woff;
                const u32 colw = (u32)(base + 32ull * wv * WPW);  // column of the wave's first word
#pragma unroll
                for (u32 m = 0; m < R; ++m) {
                    const u32 inc = dev::wave_incl_sum_dpp(c[m]);
                    const u32 rt = (u32)__builtin_amdgcn_readlane((int)inc, 63);
                    if (rt == 0) continue;  // uniform
                    u64 q = pos + (inc - c[m]);
                    const u32 cw = colw + 32u * ((m * dev::WAVE + (u32)l) * E);
                    u32 x[E];
                    if (c[m]) words(m, x);
#pragma unroll
                    for (u32 e = 0; e < E; ++e) {
                        u32 w = c[m] ? x[e] : 0u;
                        while (w) {
                            // the numeric pass's slot is exact; a count that
                            // disagrees never writes past it
                            if (MODE == ONE || q < lim) out[q] = cw + 32u * e + (u32)__builtin_ctz(w);
                            ++q;
                            w &= w - 1u;
                        }
                    }
                    pos += rt;
                }
            }
            written += tot;
        }
        __syncthreads();
    }
    if (MODE == SYM) {
        const u32 n = dev::block_sum<u32, WG>(mine, red);
        if (threadIdx.x == 0) o.nnz_row[r] = n;
    } else if (dumping) {
        // listed for emission: in the one pass always (it is written at its
        // exact offset, its capacity does not matter); in the numeric pass
        // when it matches the symbolic count
        if (threadIdx.x == 0) {
            // every claimed slot is emitted (the emission never writes past
            // the row's exact size)
            o.dump.list[slot] = idx;
            const bool fits = MODE == ONE ? written <= 0xFFFFFFFFull : written == lim;
            if (fits) {
                o.dump.isdense[r] = 1;
                if (MODE == ONE) o.nnz_row[r] = (u32)written;
            } else if (MODE == ONE) {
                o.nnz_row[r] = FAILED;
                atomicAdd(o.overflow, 1u);
            } else {
                atomicOr(o.overflow, 2u);
            }
        }
    } else if (threadIdx.x == 0) {
        if (MODE == ONE) {
            // the capacity was too small: the partial output is discarded
            o.nnz_row[r] = written <= lim ? (u32)written : FAILED;
            if (written > lim) atomicAdd(o.overflow, 1u);
        } else if (written != lim) {
            atomicOr(o.overflow, 2u);
        }
    }
}

// BITMAP rows: one workgroup per row (a launch of one workgroup per row lets
// the dispatcher balance rows of very different sizes over the CUs and
// interleave the concurrent hash launches; a persistent grid did neither).
template <int WBITS, int WG, int MODE, bool COMB>
__global__ __launch_bounds__(WG) void k_bitmap(Mats g, const u32 *__restrict__ perm, Out o) {
    __shared__ __align__(16) u32 bm[1u << WBITS];
    __shared__ __align__(8) u32 big[BIG_WORDS(WG)];
    __shared__ u32 red[WG / 64 + 1];
    bitmap_row<WBITS, WG, MODE, COMB>(g, perm, blockIdx.x, o, bm, big, red);
}

// The same rows, persistent: a grid of the workgroups that fit on the device
// at once, each taking its next row from the counter *ctr (zero at launch)
// as soon as it finishes one -- dynamic balance over rows of very different
// sizes without a dispatch per row.  The next row is claimed at the start of
// the current one.
template <int WBITS, int WG, int MODE, bool COMB>
__global__ __launch_bounds__(WG) void k_bitmap_persist(Mats g, const u32 *__restrict__ perm, u32 count,
                                                       u32 *__restrict__ ctr, Out o) {
    __shared__ __align__(16) u32 bm[1u << WBITS];
    __shared__ __align__(8) u32 big[BIG_WORDS(WG)];
    __shared__ u32 red[WG / 64 + 1];
    __shared__ u32 next;
    if (threadIdx.x == 0) next = atomicAdd(ctr, 1u);
    __syncthreads();
    u32 idx = next;
    __syncthreads();
    while (idx < count) {
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1u);
        bitmap_row<WBITS, WG, MODE, COMB>(g, perm, idx, o, bm, big, red);
        __syncthreads();
        idx = next;
        __syncthreads();
    }
}

}  // namespace spg
}  // namespace clb
