// spgemm.hip -- host driver of the boolean SpGEMM C = A * B (DCSR).
//
// Plain mode follows the two-pass structure of clBool's
// dcsr::matrix_multiplication_hash (src/dcsr/dcsr_matrix_multiplication_hash.cpp:51-119):
// count products per row (count_workload, dcsr_matrix_multiplication.cpp:388-404),
// bin rows (build_groups_and_allocate_hash, hash.cpp:301-345 -- here on the
// device, no per-row host loop), symbolic pass (count_nnz, :122-201), scan
// (fill_nnz's prefix_sum, :218), numeric pass (:235-280), drop empty rows
// (prepare_positions + set_positions, :282-296).  Learned mode lives in
// learned.hip and shares the kernels.
#include "spgemm.hpp"

#include <chrono>
#include <sstream>

namespace clb {
namespace spg {

// Mats::splits of B: for each stored row, the start of each column slice of
// SPAN columns (lower bounds in the sorted row) and its end.
__global__ void k_slice_splits(const u32 *__restrict__ b_rpt, const u32 *__restrict__ b_rows,
                               const u32 *__restrict__ b_cols, u32 b_nzr, u32 mapn, u32 nsl, u64 span,
                               u32 *__restrict__ splits) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b_nzr) return;
    const u32 k = b_rows[i];
    if (k >= mapn) return;
    const u32 beg = b_rpt[i], end = b_rpt[i + 1];
    u32 *out = splits + (u64)k * (nsl + 1);
    u32 lo = beg;
    out[0] = beg;
    for (u32 s = 1; s < nsl; ++s) {
        const u64 bound = (u64)s * span;
        u32 hi = end;
        while (lo < hi) {  // first column >= bound, searching on from the previous slice
            const u32 mid = lo + (hi - lo) / 2;
            if ((u64)b_cols[mid] < bound) lo = mid + 1;
            else hi = mid;
        }
        out[s] = lo;
    }
    out[nsl] = end;
}

__global__ void k_brange_set(const u32 *__restrict__ b_rpt, const u32 *__restrict__ b_rows, u32 b_nzr, u32 mapn,
                             uint2 *__restrict__ brange, u32 *__restrict__ bpres) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b_nzr) return;
    const u32 r = b_rows[i];
    if (r >= mapn) return;
    brange[r] = make_uint2(b_rpt[i], b_rpt[i + 1]);
    if (bpres) atomicOr(&bpres[r >> 5], 1u << (r & 31));
}

// Histogram of bins.  key == nullptr: key = min(P, ncols) (symbolic sizing).
// Rows in BIN_EMPTY get nnz_row = 0 when zero_empty is set.
__global__ __launch_bounds__(256) void k_bin_hist(Mats g, const u32 *__restrict__ P, const u32 *__restrict__ key,
                                                  u32 n, BinCfg c, u32 *__restrict__ hist, u64 *__restrict__ sums,
                                                  u32 *__restrict__ nnz_row) {
    __shared__ u32 h[NBINS];
    __shared__ unsigned long long hs[3][NBINS];  // products, A nonzeros, key (output nnz)
    if (threadIdx.x < NBINS) {
        h[threadIdx.x] = 0;
        hs[0][threadIdx.x] = hs[1][threadIdx.x] = hs[2][threadIdx.x] = 0;
    }
    __syncthreads();
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const u32 p = P[i];
        const u32 k = key ? key[i] : (p < c.ncols ? p : c.ncols);
        const u32 b = classify(p, k, c);
        atomicAdd(&h[b], 1u);
        atomicAdd(&hs[0][b], (unsigned long long)p);
        atomicAdd(&hs[1][b], (unsigned long long)(g.a_rpt[g.row0 + i + 1] - g.a_rpt[g.row0 + i]));
        if (key) atomicAdd(&hs[2][b], (unsigned long long)k);
        if (b == BIN_EMPTY && nnz_row) nnz_row[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x < NBINS && h[threadIdx.x]) {
        atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
        for (int q = 0; q < 3; ++q)
            atomicAdd(reinterpret_cast<unsigned long long *>(&sums[q * NBINS + threadIdx.x]), hs[q][threadIdx.x]);
    }
}

// Exclusive scan of the bin histogram into cursor[0..NBINS] (one wave).
__global__ void k_bin_offsets(const u32 *__restrict__ hist, u32 *__restrict__ cursor) {
    const int l = dev::lane();
    const u32 v = l < (int)NBINS ? hist[l] : 0u;
    const u32 inc = dev::wave_incl_scan(v);
    if (l < (int)NBINS) cursor[l] = inc - v;
    if (l == (int)NBINS - 1) cursor[NBINS] = inc;
}

// Rows grouped by bin (order inside a bin is not significant: output
// positions come from the scan of row sizes, not from the processing order).
__global__ __launch_bounds__(256) void k_bin_scatter(const u32 *__restrict__ P, const u32 *__restrict__ key, u32 n,
                                                     BinCfg c, u32 *__restrict__ cursor, u32 *__restrict__ perm) {
    __shared__ u32 h[NBINS];
    __shared__ u32 base[NBINS];
    if (threadIdx.x < NBINS) h[threadIdx.x] = 0;
    __syncthreads();
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    u32 b = BIN_EMPTY, local = 0;
    if (i < n) {
        const u32 p = P[i];
        const u32 k = key ? key[i] : (p < c.ncols ? p : c.ncols);
        b = classify(p, k, c);
        if (b != BIN_EMPTY) local = atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NBINS && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]);
    __syncthreads();
    if (i < n && b != BIN_EMPTY) perm[base[b] + local] = i;
}

__global__ void k_nonzero_flags(const u32 *__restrict__ nnz_row, u32 n, u32 *__restrict__ flags) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = nnz_row[i] != 0;
}

// prepare_for_shift_empty_rows + set_positions_pointers_and_rows
// (prepare_positions.cl:69-84, set_positions.cl:46-66).
__global__ void k_set_positions(const u32 *__restrict__ nnz_row, const u64 *__restrict__ off,
                                const u32 *__restrict__ pos, u32 n, const u32 *__restrict__ a_rows, u32 row0,
                                u32 *__restrict__ c_rpt, u32 *__restrict__ c_rows, u32 nnz, u32 nzr) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) c_rpt[nzr] = nnz;
    if (i >= n || nnz_row[i] == 0) return;
    const u32 q = pos[i];
    c_rpt[q] = static_cast<u32>(off[i]);
    c_rows[q] = a_rows[row0 + i];
}

// Writes one dumped row (one workgroup per row) into its columns at their
// final positions.  Each of the 2 waves owns a contiguous half of the row's
// words; its start position is the prefix of the chunk counts before its
// first chunk, so the waves need no exchange.  A bitmap slice is loaded 1024
// words at a time (four rounds of 256, 16 bytes per lane per round),
// double-buffered (the next 1024 words' loads are in flight while the
// current ones are expanded); a slice dumped as (word index, word) pairs is
// read 256 pairs per round.  Either way the words go through the wave's LDS
// staging area (dev::emit_round4_staged: one wave scan per round), so every
// store instruction is one contiguous run of columns.  Two-wave workgroups
// (8.4 KB of staging per wave) let nine rows share a CU, so one row's load
// latency overlaps another's expansion (at 256 threads, four rows: 11%
// slower).  gpre = exclusive prefix of chunk_cnt over all slots.
constexpr int EMIT_WG = 128;
__global__ __launch_bounds__(EMIT_WG) void k_bitmap_emit(const uint4 *__restrict__ store, u32 nch,
                                                         const u64 *__restrict__ gpre, const u32 *__restrict__ list,
                                                         const u32 *__restrict__ perm, const u64 *__restrict__ off,
                                                         u32 words_per_row, u32 slice_words,
                                                         const u32 *__restrict__ fmt, u32 *__restrict__ cols) {
    constexpr u32 NW = EMIT_WG / 64, ROUNDS = 4, SPAN = ROUNDS * 256;
    __shared__ u32 stage[NW][dev::STAGE_WORDS];
    const u32 slot = blockIdx.x;  // slots are claimed densely from 0
    const u32 r = perm[list[slot]];
    u32 *out = cols + off[r];
    const u64 lim = off[r + 1] - off[r];
    const u32 wv = dev::wave_id(), l = (u32)dev::lane();
    // words_per_row and slice_words are multiples of 2^14, so wpw and every
    // slice boundary inside it are multiples of 2 * SPAN and of EMIT_CHUNK
    const u32 wpw = words_per_row / NW;
    const u32 first = wv * wpw, last = first + wpw;
    const u32 nsl = words_per_row / slice_words;
    const u32 *row = reinterpret_cast<const u32 *>(store) + (u64)slot * words_per_row;
    u64 p = gpre[(u64)slot * nch + first / EMIT_CHUNK] - gpre[(u64)slot * nch];
    // a bitmap batch is ROUNDS rounds of 256 words, lane l holding words
    // 4l .. 4l + 3 of each (16-byte loads); s0 counts words from the row's start
    auto load = [&](u32(&w)[ROUNDS][4], u32 s0) {
#pragma unroll
        for (u32 k = 0; k < ROUNDS; ++k) {
            const uint4 v = *reinterpret_cast<const uint4 *>(row + s0 + k * 256 + 4 * l);
            w[k][0] = v.x;
            w[k][1] = v.y;
            w[k][2] = v.z;
            w[k][3] = v.w;
        }
    };
    auto expand = [&](const u32(&w)[ROUNDS][4], u32 s0) {
#pragma unroll
        for (u32 k = 0; k < ROUNDS; ++k) {
            const u32 c0 = (s0 + k * 256 + 4 * l) * 32u;
            const u32 cw[4] = {c0, c0 + 32u, c0 + 64u, c0 + 96u};
            p += dev::emit_round4_staged(w[k], cw, p, out, lim, stage[wv]);
        }
    };
    // the wave's words one slice at a time
    for (u32 seg = first; seg < last;) {
        const u32 sidx = seg / slice_words, s0w = sidx * slice_words;
        const u32 send = s0w + slice_words < last ? s0w + slice_words : last;
        const u32 f = fmt[(u64)slot * nsl + sidx];
        if (f) {
            // the pairs whose (slice-relative) word index is in [seg, send)
            const uint2 *pairs = reinterpret_cast<const uint2 *>(row + (u64)s0w);
            const u32 npairs = f - 1u;
            auto lower = [&](u32 key) {
                u32 lo = 0, hi = npairs;
                while (lo < hi) {
                    const u32 mid = (lo + hi) / 2;
                    if (pairs[mid].x < key) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            const u32 pa = seg == s0w ? 0u : lower(seg - s0w);
            const u32 pb = send == s0w + slice_words ? npairs : lower(send - s0w);
            for (u32 k0 = pa; k0 < pb; k0 += 256) {
                u32 x[4], cw[4];
#pragma unroll
                for (u32 e = 0; e < 4; ++e) {
                    const u32 k = k0 + 4 * l + e;
                    if (k < pb) {
                        const uint2 pr = pairs[k];
                        x[e] = pr.y;
                        cw[e] = (s0w + pr.x) * 32u;
                    } else {
                        x[e] = 0;
                        cw[e] = 0;
                    }
                }
                p += dev::emit_round4_staged(x, cw, p, out, lim, stage[wv]);
            }
        } else {
            u32 wa[ROUNDS][4], wb[ROUNDS][4];
            load(wa, seg);
            for (u32 s0 = seg; s0 < send; s0 += 2 * SPAN) {
                load(wb, s0 + SPAN);
                expand(wa, s0);
                if (s0 + 2 * SPAN < send) load(wa, s0 + 2 * SPAN);
                expand(wb, s0 + SPAN);
            }
        }
        seg = send;
    }
}

// ---------------------------------------------------------------------------
// Rows of the bitmap range of perm whose size key (capacity or exact size)
// reaches thr: the rows likely to be dumped.
__global__ void k_dump_candidates(const u32 *__restrict__ perm, u32 n, const u32 *__restrict__ key, u32 thr,
                                  u32 *__restrict__ count) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool hit = i < n && key[perm[i]] >= thr;
    const unsigned long long m = __ballot(hit);
    if (dev::lane() == 0 && m) atomicAdd(count, (u32)__popcll(m));
}

void setup_dump(clb_controls_s *ctl, const Setup &s, const BinHost &h, DumpBufs &d, Out &o, bool one_pass,
                const u32 *perm, const u32 *key) {
    o.dump = Dump{};
    d.on = false;
    // default: on in the one pass (its in-kernel emission writes a temporary
    // slot that the compaction copies again), off in the numeric pass (which
    // writes the final positions directly)
    const u32 knob = ctl->tuning.dump_bits_per_chunk ? ctl->tuning.dump_bits_per_chunk : (one_pass ? 1024u : 65536u);
    if (s.cfg.bm_bits < 14 || knob >= 65536u) return;
    const u64 nslices = ((u64)s.g.ncols + (32ull << s.cfg.bm_bits) - 1) / (32ull << s.cfg.bm_bits);
    d.nbm = h.off[NBINS] - h.off[BIN_BITMAP0];
    // a dumped row keeps every slice: nslices windows of 2^bm_bits words
    if (d.nbm == 0 || (nslices << s.cfg.bm_bits) > (1ull << 26)) return;
    d.words = (u32)(nslices << s.cfg.bm_bits);
    d.slice_words = 1u << s.cfg.bm_bits;
    d.nch = d.words / EMIT_CHUNK;
    const u32 dense_min = ((1u << s.cfg.bm_bits) / 1024u) * knob;
    // slots for the rows likely to be dumped: those whose size key reaches a
    // quarter of the dense threshold over all slices, plus a margin for rows
    // the one pass dumps because they outgrow their capacity
    u32 slots = d.nbm;
    if (key && perm) {
        DevArray<u32> cnt(ctl, 1);
        fill_u32(ctl, cnt.ptr, 1, 0u);
        const u64 thr = (u64)dense_min * nslices / 4;
        k_dump_candidates<<<div_up(d.nbm, 256), 256, 0, ctl->stream>>>(
            perm + h.off[BIN_BITMAP0], d.nbm, key, (u32)(thr < 0xFFFFFFFFull ? thr : 0xFFFFFFFFull), cnt.ptr);
        CLB_LAUNCHED("dump_candidates");
        u32 cand = 0;
        readback(ctl, &cand, cnt.ptr, 4);
        const u64 want = (u64)cand + cand / 8 + 256;
        slots = want < d.nbm ? (u32)want : d.nbm;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    // keep most of HBM for the operands and the output; blocks cached by the
    // pool count as free (the pool hands them back when an allocation fails)
    const u64 budget = (free_b + pool_for(ctl->device).cached_bytes()) / 3;
    const u64 per_slot = (u64)d.words * 4;
    if ((u64)slots * per_slot > budget) slots = (u32)(budget / per_slot);
    if (slots == 0) return;
    d.slots = slots;
    d.store.reset(ctl, (u64)slots * d.words / 4);
    d.cnt.reset(ctl, (u64)slots * d.nch);
    d.list.reset(ctl, slots);
    d.count.reset(ctl, 1);
    d.isdense.reset(ctl, s.nrows);
    d.fmt.reset(ctl, (u64)slots * nslices);  // written for every slice of a claimed slot
    fill_u32(ctl, d.cnt.ptr, d.cnt.n, 0);
    fill_u32(ctl, d.count.ptr, 1, 0);
    fill_u32(ctl, d.isdense.ptr, s.nrows, 0);
    // the decision is taken on slice 0's count
    o.dump = Dump{d.store.ptr, d.cnt.ptr, d.list.ptr, d.count.ptr, d.isdense.ptr, dense_min, slots, d.fmt.ptr};
    d.on = true;
}

u32 emit_prepare(clb_controls_s *ctl, DumpBufs &d, DevArray<u64> &gpre, clb_spgemm_stats *st) {
    if (!d.on) return 0;
    u32 count = 0;
    readback(ctl, &count, d.count.ptr, 4);
    if (count > d.slots) count = d.slots;  // claims beyond the last slot were refused
    if (!count) return 0;
    if (d.words % (1u << 14)) raise(RUN_KERNEL_FAILURE, "dumped rows must hold whole 2^14-word windows");
    gpre.reset(ctl, d.cnt.n + 1);
    exclusive_scan_u32_to_u64(ctl, d.cnt.ptr, gpre.ptr, d.cnt.n, gpre.ptr + d.cnt.n);
    if (st) {
        u64 total = 0;
        readback(ctl, &total, gpre.ptr + d.cnt.n, 8);
        st->dense_rows += count;
        st->dense_nnz += total;
    }
    return count;
}

void emit_launch(hipStream_t es, u32 count, const u32 *perm, const BinHost &h, const DumpBufs &d, const u64 *gpre,
                 const u64 *off, u32 *cols, LaunchTimer *lt, int pass) {
    if (!count) return;
    if (lt) lt->begin(es);
    k_bitmap_emit<<<count, EMIT_WG, 0, es>>>(d.store.ptr, d.nch, gpre, d.list.ptr, perm + h.off[BIN_BITMAP0], off,
                                            d.words, d.slice_words, d.fmt.ptr, cols);
    CLB_LAUNCHED("bitmap_emit");
    if (lt) lt->end((int)CLB_KF(pass, CLB_KK_EMIT), es);
}

void emit_dense(clb_controls_s *ctl, const u32 *perm, const BinHost &h, DumpBufs &d, const u64 *off, u32 *cols,
                LaunchTimer *lt, int pass, clb_spgemm_stats *st) {
    DevArray<u64> gpre;  // released after the launch: later work on this stream is ordered after it
    const u32 count = emit_prepare(ctl, d, gpre, st);
    emit_launch(ctl->stream, count, perm, h, d, gpre.ptr, off, cols, lt, pass);
}

BinCfg make_bincfg(const clb_controls_s *ctl, u32 ncols) {
    BinCfg c{};
    c.ncols = ncols;
    c.tiny_max = ctl->tuning.tiny_max_products;
    u32 hb = 0;
    while ((1u << (hb + 1)) <= ctl->tuning.hash_max_slots) ++hb;
    c.hash_max_bits = hb < HASH_MIN_BITS ? HASH_MIN_BITS : (hb > HASH_MAX_BITS ? HASH_MAX_BITS : hb);
    c.bitmap_min = ctl->tuning.bitmap_min_products;
    c.bitmap_ratio = ctl->tuning.bitmap_slot_ratio;
    // Window of a bitmap kernel: the matrix width rounded up to a power of
    // two words, at most 2^15 words (128 KB of LDS, one 1024-thread workgroup
    // per CU); wider matrices are processed in column slices.  The kernels
    // that also write columns use 2^14 words (64 KB, two 512-thread
    // workgroups per CU, so one row's emission overlaps the other's gather)
    // when that costs at most one extra slice, i.e. up to 2^20 columns.  On
    // wider matrices each extra slice re-reads the A row and its B-row
    // pieces; whether the overlap still pays depends on the rows (with
    // dense-row dumps and the ESC bin it does at R-MAT scale 22), so the
    // bitmap_window_bits knob can force the window.
    const u64 words = ((u64)ncols + 31) / 32;
    u32 wb = 6;
    while (wb < 15 && (1ull << wb) < words) ++wb;
    c.sym_bits = wb;
    c.bm_bits = (wb == 15 && words <= (1ull << 15)) ? 14 : wb;
    const u32 forced = ctl->tuning.reserved[CLB_TUNE_BITMAP_WINDOW_BITS];
    if (forced >= 13 && forced <= 15) c.bm_bits = forced < wb ? forced : wb;
    c.multi_slice = (u64)ncols > (32ull << c.bm_bits) ? 1u : 0u;
    c.tables_over_slices = (ctl->tuning.reserved[CLB_TUNE_HASH_WIDE] >> CLB_TUNE_TABLES_OVER_SLICES_BIT) & 1u;
    const u32 em = ctl->tuning.reserved[CLB_TUNE_ESC_MAX_PRODUCTS];
    c.esc_max = em < (1u << ESC_MAX_BITS) ? em : (1u << ESC_MAX_BITS);
    c.esc_only = 0;
    return c;
}

// The per-slice split table of B (Mats::splits) for bitmap rows over
// several column windows: built on the first pass that has bitmap rows.
void ensure_splits(clb_controls_s *ctl, Setup &s, const BinHost &h) {
    if (s.g.splits || h.off[NBINS] == h.off[BIN_BITMAP0]) return;
    const clb_dcsr &b = s.b;
    const u64 span = 32ull << s.cfg.bm_bits;
    const u64 nsl = ((u64)b.ncols + span - 1) / span;
    if (nsl <= 1 || !b.nzr || (u64)s.mapn * (nsl + 1) >= (1ull << 32)) return;
    // zero for the rows B does not store (an empty part in every slice), so
    // the gather reads it without looking the row up in brange
    s.splits.reset(ctl, (u64)s.mapn * (nsl + 1));
    fill_u32(ctl, s.splits.ptr, s.splits.n, 0u);
    k_slice_splits<<<div_up(b.nzr, 256), 256, 0, ctl->stream>>>(u32p(b.rpt), u32p(b.rows), u32p(b.cols), b.nzr,
                                                                 s.mapn, (u32)nsl, span, s.splits.ptr);
    CLB_LAUNCHED("slice_splits");
    s.g.splits = s.splits.ptr;
    s.g.nsl = (u32)nsl;
}

void prepare(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 nrows, Setup &s,
             bool count, const Setup *shared) {
    s.nrows = nrows;
    s.row0 = row_begin;
    if (shared) {
        // a plan's block: B's row map and split table were built once
        s.mapn = shared->mapn;
        s.g = shared->g;
        s.cfg = shared->cfg;
        s.b = shared->b;
    } else {
        s.mapn = b.nrows;
        s.brange.reset(ctl, s.mapn ? s.mapn : 1);
        // hypersparse B (under one stored row in 16): a presence bit per row id
        // instead of clearing the whole row map
        const bool sparse_map = (u64)b.nzr * 16 < s.mapn;
        if (sparse_map) {
            s.bpres.reset(ctl, ((u64)s.mapn + 31) / 32);
            fill_u32(ctl, s.bpres.ptr, s.bpres.n, 0u);
        } else {
            fill_u64(ctl, reinterpret_cast<u64 *>(s.brange.ptr), s.mapn, 0ull);
        }
        if (b.nzr) {
            k_brange_set<<<div_up(b.nzr, 256), 256, 0, ctl->stream>>>(u32p(b.rpt), u32p(b.rows), b.nzr, s.mapn,
                                                                       s.brange.ptr,
                                                                       sparse_map ? s.bpres.ptr : nullptr);
            CLB_LAUNCHED("brange_set");
        }
        s.g.bpres = sparse_map ? s.bpres.ptr : nullptr;
        s.g.a_rpt = u32p(a.rpt);
        s.g.a_rows = u32p(a.rows);
        s.g.a_cols = u32p(a.cols);
        s.g.brange = s.brange.ptr;
        s.g.b_cols = u32p(b.cols);
        s.g.mapn = s.mapn;
        s.g.ncols = b.ncols;
        s.g.emit = ctl->tuning.emit_lane_bits_min;
        const u32 sb = ctl->tuning.reserved[CLB_TUNE_EMIT_SPARSE_BITS];
        s.g.sparse_bits = sb ? sb : 256u;
        s.g.combine = ctl->tuning.reserved[CLB_TUNE_GATHER_COMBINE];
        s.cfg = make_bincfg(ctl, b.ncols);
        s.g.splits = nullptr;
        s.g.nsl = 0;
        s.g.split_bits = s.cfg.bm_bits;
        s.b = b;  // the split table is built by ensure_splits, once bitmap rows exist
    }
    s.g.row0 = row_begin;
    s.prod.reset(ctl, nrows);
    s.P = s.prod.ptr;
    s.ctr.reset(ctl, 4);
    s.scal.reset(ctl, 8);  // [0] total products, [1] nnz total, [2] nzr total, [3] overflow flag
    fill_u64(ctl, s.scal.ptr, 8, 0ull);
    if (!count) return;
    // count_workload (count_workload.cl:30-53): products of each A row,
    // balanced over A's nonzeros (k_row_stats)
    fill_u32(ctl, s.prod.ptr, nrows, 0u);
    k_row_stats<false><<<div_up(a.nnz, RS_WG), RS_WG, 0, ctl->stream>>>(s.g, nrows, s.prod.ptr, s.scal.ptr, nullptr,
                                                                        nullptr, nullptr, nullptr);
    CLB_LAUNCHED("count_workload");
}

void bin_rows(clb_controls_s *ctl, const Setup &s, const u32 *key, u32 *nnz_row_zero, DevArray<u32> &perm,
              BinBuf &bb) {
    bb.buf.reset(ctl, 2 * NBINS + 2 + 2 * 3 * NBINS);
    fill_u32(ctl, bb.buf.ptr, bb.buf.n, 0u);
    perm.reset(ctl, s.nrows ? s.nrows : 1);
    const u32 grid = div_up(s.nrows, 256);
    k_bin_hist<<<grid, 256, 0, ctl->stream>>>(s.g, s.P, key, s.nrows, s.cfg, bb.hist(), bb.sums(),
                                              nnz_row_zero);
    CLB_LAUNCHED("bin_hist");
    k_bin_offsets<<<1, 64, 0, ctl->stream>>>(bb.hist(), bb.cursor());
    CLB_LAUNCHED("bin_offsets");
    k_bin_scatter<<<grid, 256, 0, ctl->stream>>>(s.P, key, s.nrows, s.cfg, bb.cursor(), perm.ptr);
    CLB_LAUNCHED("bin_scatter");
}

void read_bins(clb_controls_s *ctl, const BinBuf &bb, BinHost &h) {
    u32 hist[NBINS];
    readback(ctl, hist, bb.hist(), sizeof(hist));
    h.off[0] = 0;
    for (u32 b = 0; b < NBINS; ++b) h.off[b + 1] = h.off[b] + hist[b];
    u64 sums[3 * NBINS];
    readback(ctl, sums, bb.sums(), sizeof(sums));
    for (u32 b = 0; b < NBINS; ++b) {
        h.products[b] = sums[b];
        h.a_nnz[b] = sums[NBINS + b];
        h.out_nnz[b] = sums[2 * NBINS + b];
    }
}

// Sums nnz_row over the rows of bins 1..NBINS-1 (perm[off[1] .. off[NBINS]));
// the EMPTY bin's range of perm is never written by the scatter.
__global__ __launch_bounds__(256) void k_sum_by_bin(const u32 *__restrict__ perm, const u32 *__restrict__ off,
                                                    u32 nrows, const u32 *__restrict__ nnz_row,
                                                    u64 *__restrict__ sums) {
    __shared__ unsigned long long hs[NBINS];
    __shared__ u32 so[NBINS + 1];
    if (threadIdx.x < NBINS) hs[threadIdx.x] = 0;
    if (threadIdx.x <= NBINS) so[threadIdx.x] = off[threadIdx.x];
    __syncthreads();
    const u32 i = so[1] + blockIdx.x * 256 + threadIdx.x;
    if (i < so[NBINS]) {
        u32 b = 1;
        while (b + 1 < NBINS && so[b + 1] <= i) ++b;
        const u32 row = perm[i];
        const u32 v = row < nrows ? nnz_row[row] : 0u;
        if (v != FAILED) atomicAdd(&hs[b], (unsigned long long)v);
    }
    __syncthreads();
    if (threadIdx.x < NBINS && hs[threadIdx.x])
        atomicAdd(reinterpret_cast<unsigned long long *>(&sums[threadIdx.x]), hs[threadIdx.x]);
}

void sum_by_bin(clb_controls_s *ctl, const u32 *perm, const BinHost &h, u32 nrows, const u32 *nnz_row,
                u64 out[NBINS]) {
    const u32 n = h.off[NBINS] - h.off[1];
    DevArray<u32> off(ctl, NBINS + 1);
    DevArray<u64> sums(ctl, NBINS);
    CLB_HIP(hipMemcpyAsync(off.ptr, h.off, 4 * (NBINS + 1), hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
    fill_u64(ctl, sums.ptr, NBINS, 0ull);
    if (n) {
        k_sum_by_bin<<<div_up(n, 256), 256, 0, ctl->stream>>>(perm, off.ptr, nrows, nnz_row, sums.ptr);
        CLB_LAUNCHED("sum_by_bin");
    }
    readback(ctl, out, sums.ptr, NBINS * 8);
}

void LaunchTimer::begin(hipStream_t st) {
    if (!on) return;
    CLB_HIP(hipEventCreate(&pending), EVENT_WAITING_ERROR);
    CLB_HIP(hipEventRecord(pending, st), EVENT_WAITING_ERROR);
}

void LaunchTimer::end(int family, hipStream_t st) {
    if (!on) return;
    hipEvent_t e;
    CLB_HIP(hipEventCreate(&e), EVENT_WAITING_ERROR);
    CLB_HIP(hipEventRecord(e, st), EVENT_WAITING_ERROR);
    recs.push_back({family, pending, e});
    pending = nullptr;
}

void LaunchTimer::collect(clb_spgemm_stats &st) {
    for (auto &r : recs) {
        float m = 0;
        if (hipEventElapsedTime(&m, r.a, r.b) == hipSuccess) {
            st.ms_kernel[r.family] += m;
            st.launches[r.family] += 1;
        }
    }
}

LaunchTimer::~LaunchTimer() {
    for (auto &r : recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    if (pending) (void)hipEventDestroy(pending);
}

// Launch the per-bin kernels of one pass.  off[b]..off[b+1] is bin b's range
// of perm.  numeric == false: write nnz_row; true: write columns.
// One workgroup per row, or (tuning.bitmap_persistent) a persistent grid
// that takes rows from a counter.
template <int WBITS, int WG, int MODE, bool COMB>
static void launch_bitmap_c(clb_controls_s *ctl, hipStream_t st, const Setup &s, const u32 *perm, u32 n,
                            const Out &o) {
    if (!ctl->tuning.bitmap_persistent) {
        k_bitmap<WBITS, WG, MODE, COMB><<<n, WG, 0, st>>>(s.g, perm, o);
        return;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bitmap_persist<WBITS, WG, MODE, COMB>, WG, 0) !=
            hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const u64 fit = (u64)ctl->cu_count * (u64)per_cu;
    u32 *ctr = s.ctr.ptr + MODE;  // one counter per pass of the SpGEMM
    fill_u32_on(st, ctr, 1, 0u);
    k_bitmap_persist<WBITS, WG, MODE, COMB><<<(u32)(n < fit ? n : fit), WG, 0, st>>>(s.g, perm, n, ctr, o);
}

template <int WBITS, int WG, int MODE>
static void launch_bitmap(clb_controls_s *ctl, hipStream_t st, const Setup &s, const u32 *perm, u32 n,
                          const Out &o) {
    if (s.g.combine) launch_bitmap_c<WBITS, WG, MODE, true>(ctl, st, s, perm, n, o);
    else launch_bitmap_c<WBITS, WG, MODE, false>(ctl, st, s, perm, n, o);
}

void join_deferred(clb_controls_s *ctl, const hipEvent_t *defer) {
    for (int k = 0; k < clb_controls_s::NSIDE; ++k)
        CLB_HIP(hipStreamWaitEvent(ctl->stream, defer[k], 0), EVENT_WAITING_ERROR);
}

void launch_bins(clb_controls_s *ctl, const Setup &s, const u32 *perm, const u32 off[NBINS + 1], int mode,
                 const Out &o, LaunchTimer *lt, int family, hipEvent_t *defer, bool one_stream) {
    if (!one_stream) fork_side(ctl);
    // table kernels alternate over side streams 1.. (the first one takes
    // stream 0 when no bitmap kernel was launched there)
    int k = 0;
    auto next = [&]() {
        constexpr int NS = clb_controls_s::NSIDE;
        if (one_stream) return ctl->stream;
        return k == 0 ? side_stream(ctl, k++) : side_stream(ctl, 1 + (k++ - 1) % (NS - 1));
    };
    LaunchTimer dummy;
    LaunchTimer &t = lt ? *lt : dummy;
    // family = the pass (CLB_KP_*); < 0 picks symbolic / numeric from the mode
    const int pass = family >= 0 ? family : (mode == SYM ? (int)CLB_KP_SYMBOLIC : (int)CLB_KP_NUMERIC);
    auto fam = [&](int kind) { return (int)CLB_KF(pass, kind); };
#define DISPATCH3(KERNEL, GRID, WG, ST, ...)                                                       \
    do {                                                                                          \
        if (mode == SYM) KERNEL<__VA_ARGS__, SYM><<<GRID, WG, 0, ST>>>(s.g, pp, o);               \
        else if (mode == NUM) KERNEL<__VA_ARGS__, NUM><<<GRID, WG, 0, ST>>>(s.g, pp, o);          \
        else KERNEL<__VA_ARGS__, ONE><<<GRID, WG, 0, ST>>>(s.g, pp, o);                           \
    } while (0)
    // BITMAP (all classes in one launch, largest rows first; the ESC SpGEMM
    // sorts these rows in global memory instead).  The longest kernel of the
    // pass goes first and on a side stream of its own, so that it never
    // queues behind a table kernel; the table kernels fill the CUs around it.
    if (!s.cfg.esc_only) {
        const u32 n = off[NBINS] - off[BIN_BITMAP0];
        if (n) {
            const u32 *pp = perm + off[BIN_BITMAP0];
            hipStream_t st = one_stream ? ctl->stream : side_stream(ctl, 0);
            k = 1;
            t.begin(st);
#define BITMAP3(WB, WGS)                                                   \
    do {                                                                   \
        if (mode == SYM) launch_bitmap<WB, WGS, SYM>(ctl, st, s, pp, n, o); \
        else if (mode == NUM) launch_bitmap<WB, WGS, NUM>(ctl, st, s, pp, n, o); \
        else launch_bitmap<WB, WGS, ONE>(ctl, st, s, pp, n, o);          \
    } while (0)
            switch (mode == SYM ? s.cfg.sym_bits : s.cfg.bm_bits) {
                case 6: BITMAP3(6, 64); break;
                case 7: BITMAP3(7, 64); break;
                case 8: BITMAP3(8, 64); break;
                case 9: BITMAP3(9, 256); break;
                case 10: BITMAP3(10, 256); break;
                case 11: BITMAP3(11, 256); break;
                case 12: BITMAP3(12, 512); break;
                case 13: BITMAP3(13, 512); break;
                case 14: BITMAP3(14, 512); break;
                case 15: BITMAP3(15, 1024); break;
            }
#undef BITMAP3
            CLB_LAUNCHED("bitmap");
            t.end(fam(2), st);
        }
    }
    // TINY
    {
        const u32 n = off[BIN_TINY + 1] - off[BIN_TINY];
        if (n) {
            hipStream_t st = next();
            const u32 *pp = perm + off[BIN_TINY];
            t.begin(st);
            if (mode == SYM) k_tiny<4, SYM><<<div_up(n, 4), 256, 0, st>>>(s.g, pp, n, o);
            else if (mode == NUM) k_tiny<4, NUM><<<div_up(n, 4), 256, 0, st>>>(s.g, pp, n, o);
            else k_tiny<4, ONE><<<div_up(n, 4), 256, 0, st>>>(s.g, pp, n, o);
            CLB_LAUNCHED("tiny");
            t.end(fam(0), st);
        }
    }
    // HASH
    for (u32 bits = HASH_MIN_BITS; bits <= HASH_MAX_BITS; ++bits) {
        const u32 b = BIN_HASH0 + (bits - HASH_MIN_BITS);
        const u32 n = off[b + 1] - off[b];
        if (!n) continue;
        const u32 *pp = perm + off[b];
        hipStream_t st = next();
        t.begin(st);
        // workgroup shape per table size: the default, or twice the threads
        // when the table size's bit of the hash_wide knob is set
        const bool wide = (ctl->tuning.reserved[CLB_TUNE_HASH_WIDE] >> (bits - HASH_MIN_BITS)) & 1u;
        switch (bits * 2 + (wide ? 1 : 0)) {
            case 14: DISPATCH3(k_hash, n, 64, st, 7, 64); break;
            case 15: DISPATCH3(k_hash, n, 128, st, 7, 128); break;
            case 16: DISPATCH3(k_hash, n, 64, st, 8, 64); break;
            case 17: DISPATCH3(k_hash, n, 128, st, 8, 128); break;
            case 18: DISPATCH3(k_hash, n, 64, st, 9, 64); break;
            case 19: DISPATCH3(k_hash, n, 128, st, 9, 128); break;
            case 20: DISPATCH3(k_hash, n, 128, st, 10, 128); break;
            case 21: DISPATCH3(k_hash, n, 256, st, 10, 256); break;
            case 22: DISPATCH3(k_hash, n, 128, st, 11, 128); break;
            case 23: DISPATCH3(k_hash, n, 256, st, 11, 256); break;
            case 24: DISPATCH3(k_hash, n, 256, st, 12, 256); break;
            case 25: DISPATCH3(k_hash, n, 512, st, 12, 512); break;
            case 26: DISPATCH3(k_hash, n, 512, st, 13, 512); break;
            case 27: DISPATCH3(k_hash, n, 1024, st, 13, 1024); break;
            case 28:
            case 29: DISPATCH3(k_hash, n, 1024, st, 14, 1024); break;
        }
        CLB_LAUNCHED("hash");
        t.end(fam(1), st);
    }
    // ESC (rows of at most 2^11 / 2^12 / 2^13 products)
    for (u32 bits = ESC_MIN_BITS; bits <= ESC_MAX_BITS; ++bits) {
        const u32 b = BIN_ESC0 + (bits - ESC_MIN_BITS);
        const u32 n = off[b + 1] - off[b];
        if (!n) continue;
        const u32 *pp = perm + off[b];
        hipStream_t st = next();
        t.begin(st);
        switch (bits) {
            case 11: DISPATCH3(k_esc, n, 256, st, 11, 256); break;
            case 12: DISPATCH3(k_esc, n, 512, st, 12, 512); break;
            case 13: DISPATCH3(k_esc, n, 1024, st, 13, 1024); break;
        }
        CLB_LAUNCHED("esc");
        t.end(fam(1), st);
    }
#undef DISPATCH3
    if (one_stream) return;
    if (!defer) {
        join_side(ctl);
        return;
    }
    for (int j = 0; j < clb_controls_s::NSIDE; ++j)
        CLB_HIP(hipEventRecord(defer[j], ctl->side[j]), EVENT_WAITING_ERROR);
}

struct Timer {
    bool on;
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    Timer(bool enable, hipStream_t s) : on(enable), st(s) {}
    void mark() {
        if (!on) return;
        hipEvent_t e;
        CLB_HIP(hipEventCreate(&e), EVENT_WAITING_ERROR);
        CLB_HIP(hipEventRecord(e, st), EVENT_WAITING_ERROR);
        ev.push_back(e);
    }
    double ms(size_t i, size_t j) const {
        if (!on || j >= ev.size()) return 0;
        float m = 0;
        (void)hipEventElapsedTime(&m, ev[i], ev[j]);
        return m;
    }
    ~Timer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

void assemble(clb_controls_s *ctl, const Setup &s, const clb_dcsr &a, const u32 *nnz_row, const u64 *off,
              const u32 *pos, u64 nnz, u32 nzr, clb_buffer_s *cols, clb_dcsr &out) {
    out.cols = cols;
    out.rpt = buffer_new(ctl, 4ull * (nzr + 1));
    out.rows = buffer_new(ctl, 4ull * (nzr ? nzr : 1));
    k_set_positions<<<div_up(s.nrows ? s.nrows : 1, 256), 256, 0, ctl->stream>>>(
        nnz_row, off, pos, s.nrows, s.g.a_rows, s.row0, u32p(out.rpt), u32p(out.rows), (u32)nnz, nzr);
    CLB_LAUNCHED("set_positions");
    out.nnz = (u32)nnz;
    out.nzr = nzr;
}

// Offsets and empty-row compaction of the exact row sizes.  Results land in
// s.scal[1] (nnz) and s.scal[2] (nzr).
void scan_sizes(clb_controls_s *ctl, Setup &s, const u32 *nnz_row, DevArray<u64> &off, DevArray<u32> &pos) {
    off.reset(ctl, s.nrows + 1);
    pos.reset(ctl, s.nrows + 1);
    // off[nrows] holds the total too (row r's size is off[r + 1] - off[r])
    exclusive_scan_u32_to_u64(ctl, nnz_row, off.ptr, s.nrows, off.ptr + s.nrows);
    CLB_HIP(hipMemcpyAsync(s.scal.ptr + 1, off.ptr + s.nrows, 8, hipMemcpyDeviceToDevice, ctl->stream),
            COPY_BUFFER_ERROR);
    k_nonzero_flags<<<div_up(s.nrows, 256), 256, 0, ctl->stream>>>(nnz_row, s.nrows, pos.ptr);
    CLB_LAUNCHED("nonzero_flags");
    exclusive_scan_u32(ctl, pos.ptr, pos.ptr, s.nrows, reinterpret_cast<u32 *>(s.scal.ptr + 2));
}

void spgemm_plain(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                  bool timing, clb_dcsr &out, clb_spgemm_stats &st, const PlanRows *pr) {
    Timer tm(timing, ctl->stream);
    tm.mark();  // 0
    const u32 nrows = row_end - row_begin;
    Setup s;
    prepare(ctl, a, b, row_begin, nrows, s, pr == nullptr, pr ? pr->shared : nullptr);
    LaunchTimer lt;
    lt.on = timing;
    st.rows = nrows;
    st.n_bins = NBINS;
    DevArray<u32> nnz_row_own;
    u32 *nnz_row = nullptr;
    if (pr) {
        // planned block: products and exact row sizes come from the plan
        s.P = pr->prod;
        nnz_row = pr->size;
        st.products = pr->products;
        tm.mark();  // 1
        tm.mark();  // 2
    } else {
        DevArray<u32> perm;
        BinBuf bb;
        BinHost h1;
        nnz_row_own.reset(ctl, nrows);
        nnz_row = nnz_row_own.ptr;
        bin_rows(ctl, s, nullptr, nnz_row, perm, bb);
        read_bins(ctl, bb, h1);
        u64 total_products = 0;
        readback(ctl, &total_products, s.scal.ptr, 8);
        st.products = total_products;
        tm.mark();  // 1
        if (total_products == 0) {
            out.nrows = a.nrows;
            out.ncols = b.ncols;
            return;
        }
        Out o1{nnz_row, nullptr, nullptr, nullptr, reinterpret_cast<u32 *>(s.scal.ptr + 3), Dump{}, 0};
        ensure_splits(ctl, s, h1);
        launch_bins(ctl, s, perm.ptr, h1.off, SYM, o1, &lt, -1);
        tm.mark();  // 2
    }
    if (st.products == 0) {
        out.nrows = a.nrows;
        out.ncols = b.ncols;
        return;
    }
    DevArray<u64> offs;
    DevArray<u32> pos;
    scan_sizes(ctl, s, nnz_row, offs, pos);
    DevArray<u32> perm2;
    BinBuf bb2;
    BinHost h2;
    bin_rows(ctl, s, nnz_row, nullptr, perm2, bb2);
    read_bins(ctl, bb2, h2);
    for (u32 k = 0; k < NBINS; ++k) {
        st.bin_rows[k] = h2.off[k + 1] - h2.off[k];
        st.bin_products[k] = h2.products[k];
        st.bin_a_nnz[k] = h2.a_nnz[k];
        st.bin_out_nnz[k] = h2.out_nnz[k];
    }
    u64 sc[4];
    readback(ctl, sc, s.scal.ptr, sizeof(sc));
    tm.mark();  // 3
    if (sc[3]) raise(RUN_KERNEL_FAILURE, "hash table overflow in the symbolic pass");
    const u64 nnz = sc[1];
    const u32 nzr = static_cast<u32>(sc[2]);
    if (nnz > 0xFFFFFFFFull) {
        std::ostringstream os;
        os << "result has " << nnz << " nonzeros, more than uint32 indices allow; multiply by row blocks";
        raise(INVALID_VALUE, os.str());
    }
    out.nrows = a.nrows;
    out.ncols = b.ncols;
    if (nnz == 0) return;
    clb_buffer_s *cols = buffer_new(ctl, 4ull * nnz);
    try {
        Out o2{nnz_row, offs.ptr, u32p(cols), nullptr, reinterpret_cast<u32 *>(s.scal.ptr + 3), Dump{}, 0};
        DumpBufs dump;
        setup_dump(ctl, s, h2, dump, o2, false, perm2.ptr, nnz_row);
        ensure_splits(ctl, s, h2);
        launch_bins(ctl, s, perm2.ptr, h2.off, NUM, o2, &lt, -1);
        emit_dense(ctl, perm2.ptr, h2, dump, offs.ptr, u32p(cols), &lt, CLB_KP_NUMERIC, timing ? &st : nullptr);
        tm.mark();  // 4
        assemble(ctl, s, a, nnz_row, offs.ptr, pos.ptr, nnz, nzr, cols, out);
        tm.mark();  // 5
        u64 flag = 0;
        readback(ctl, &flag, s.scal.ptr + 3, 8);
        if (flag) raise(RUN_KERNEL_FAILURE, "numeric pass disagreed with the symbolic row sizes");
    } catch (...) {
        (void)hipStreamSynchronize(ctl->stream);
        clb_dcsr_release(&out);
        if (!out.cols) buffer_release(cols);
        throw;
    }
    st.nnz = nnz;
    st.nzr = nzr;
    lt.collect(st);
    if (timing) {
        st.ms_setup = tm.ms(0, 1);
        st.ms_symbolic = tm.ms(1, 2);
        st.ms_numeric = tm.ms(3, 4);
        st.ms_assemble = tm.ms(2, 3) + tm.ms(4, 5);
        st.ms_total = tm.ms(0, 5);
    }
}

void plan_sizes_plain(clb_controls_s *ctl, Setup &s, u32 *size) {
    // the symbolic pass of spgemm_plain over every stored row of A
    DevArray<u32> perm;
    BinBuf bb;
    BinHost h;
    bin_rows(ctl, s, nullptr, size, perm, bb);
    read_bins(ctl, bb, h);
    Out o{size, nullptr, nullptr, nullptr, reinterpret_cast<u32 *>(s.scal.ptr + 3), Dump{}, 0};
    ensure_splits(ctl, s, h);
    launch_bins(ctl, s, perm.ptr, h.off, SYM, o, nullptr, -1);
    u64 flag = 0;
    readback(ctl, &flag, s.scal.ptr + 3, 8);
    if (flag) raise(RUN_KERNEL_FAILURE, "hash table overflow in the symbolic pass");
}

void check_dims(const clb_dcsr &a, const clb_dcsr &b) {
    // hash.cpp:56-61
    if (a.ncols != b.nrows) {
        std::ostringstream s;
        s << "Invalid input matrix size! a : " << a.nrows << " x " << a.ncols << ", b: " << b.nrows << " x "
          << b.ncols;
        raise(INVALID_ARGUMENT, s.str());
    }
}

void spgemm_dispatch(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, const clb_spgemm_opts &o,
                     clb_dcsr &out, clb_spgemm_stats &st) {
    std::memset(&st, 0, sizeof(st));
    check_dims(a, b);
    clb_dcsr r{};
    r.nrows = a.nrows;
    r.ncols = b.ncols;
    u32 rb = o.row_begin, re = o.row_end ? o.row_end : a.nzr;
    if (re > a.nzr || rb > re) raise(INVALID_ARGUMENT, "row block outside the stored rows of A");
    // hash.cpp:64-67: an empty operand gives an empty result
    if (a.nnz == 0 || b.nnz == 0 || rb == re) {
        out = r;
        return;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (o.mode == CLB_SPGEMM_LEARNED || o.mode == CLB_SPGEMM_BOUND)
        spgemm_learned(ctl, a, b, rb, re, o.timing != 0, r, st, o.mode == CLB_SPGEMM_BOUND);
    else if (o.mode == CLB_SPGEMM_PLAIN)
        spgemm_plain(ctl, a, b, rb, re, o.timing != 0, r, st);
    else
        raise(INVALID_ARGUMENT, "unknown SpGEMM mode");
    if (!o.timing)
        st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctl->last = st;
    out = r;
}

}  // namespace spg
}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_dcsr_spgemm_hash(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out) {
    return run_guarded([&] {
        clb_spgemm_opts o{};
        o.mode = CLB_SPGEMM_PLAIN;
        clb_spgemm_stats st;
        clb_dcsr r{};
        spg::spgemm_dispatch(ctl, *a, *b, o, r, st);
        *out = r;
    });
}

clb_status clb_dcsr_spgemm(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, const clb_spgemm_opts *opts,
                           clb_dcsr *out, clb_spgemm_stats *stats) {
    return run_guarded([&] {
        clb_spgemm_opts o{};
        if (opts) o = *opts;
        clb_spgemm_stats st;
        clb_dcsr r{};
        spg::spgemm_dispatch(ctl, *a, *b, o, r, st);
        *out = r;
        if (stats) *stats = st;
    });
}

}  // extern "C"
