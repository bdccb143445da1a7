// sort.hip -- device sort and deduplication of (row, col) pairs.
//
// Replaces coo::sort_arrays (src/coo/coo_initialization.cpp:6-32, a bitonic
// network over round_to_power2(n) padded pairs, src/cl/coo_bitonic_sort.cl)
// and reduce_duplicates2 (src/core/matrix_coo.cpp:117-170).  Pairs are packed
// into 64-bit keys row << 32 | col and sorted by a stable LSD radix sort over
// only the bits the matrix shape needs (8-bit digits; per pass one histogram
// kernel, one scan, one scatter whose in-tile ranks come from wave ballots).
// Deduplication is flag, scan, scatter.
#include "sort.hpp"

#include "device_utils.hpp"

namespace clb {

constexpr int RS_WG = 256;
constexpr int RS_ITEMS = 16;
constexpr u32 RS_TILE = RS_WG * RS_ITEMS;
constexpr int RS_RADIX = 256;

__global__ void k_pack_pairs(const u32 *__restrict__ rows, const u32 *__restrict__ cols, u64 n,
                             u64 *__restrict__ keys) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ((u64)rows[i] << 32) | cols[i];
}

__global__ void k_unpack_pairs(const u64 *__restrict__ keys, u64 n, u32 *__restrict__ rows, u32 *__restrict__ cols) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const u64 k = keys[i];
        rows[i] = (u32)(k >> 32);
        cols[i] = (u32)k;
    }
}

// Per-tile digit histogram, stored digit-major: counts[d * tiles + tile].
__global__ __launch_bounds__(RS_WG) void k_radix_hist(const u64 *__restrict__ keys, u64 n, int shift, u32 tiles,
                                                     u32 *__restrict__ counts) {
    __shared__ u32 h[RS_RADIX];
    for (int d = threadIdx.x; d < RS_RADIX; d += RS_WG) h[d] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * RS_TILE;
#pragma unroll
    for (int q = 0; q < RS_ITEMS; ++q) {
        const u64 i = base + (u64)q * RS_WG + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RS_RADIX; d += RS_WG) counts[(u64)d * tiles + blockIdx.x] = h[d];
}

// Stable scatter.  Keys of a tile are visited in index order in rounds of
// RS_WG (one per thread); a key's rank among equal digits of its wave comes
// from eight ballots, waves are combined through LDS in wave order.
__global__ __launch_bounds__(RS_WG) void k_radix_scatter(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                                                        int shift, u32 tiles, const u32 *__restrict__ offsets) {
    constexpr int NW = RS_WG / 64;
    __shared__ u32 base[RS_RADIX];
    __shared__ u32 wcnt[NW][RS_RADIX];
    const int l = dev::lane(), w = dev::wave_id();
    for (int d = threadIdx.x; d < RS_RADIX; d += RS_WG) base[d] = offsets[(u64)d * tiles + blockIdx.x];
    const u64 tile0 = (u64)blockIdx.x * RS_TILE;
    const unsigned long long lt = (1ull << l) - 1ull;
    for (int q = 0; q < RS_ITEMS; ++q) {
        for (int d = threadIdx.x; d < NW * RS_RADIX; d += RS_WG) (&wcnt[0][0])[d] = 0;
        __syncthreads();
        const u64 i = tile0 + (u64)q * RS_WG + threadIdx.x;
        const bool valid = i < n;
        const u64 key = valid ? in[i] : 0;
        const u32 d = valid ? (u32)((key >> shift) & 0xFF) : 0;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long m = __ballot(((d >> b) & 1u) != 0);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const u32 rank = (u32)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (u32)__popcll(peers);
        __syncthreads();
        u32 before = 0;
        for (int v = 0; v < w; ++v) before += wcnt[v][d];
        if (valid) out[base[d] + before + rank] = key;
        __syncthreads();
        for (int dd = threadIdx.x; dd < RS_RADIX; dd += RS_WG) {
            u32 s = 0;
            for (int v = 0; v < NW; ++v) s += wcnt[v][dd];
            base[dd] += s;
        }
        __syncthreads();
    }
}

static int bits_for(u64 maxv) {
    int b = 0;
    while (b < 32 && (1ull << b) <= maxv) ++b;
    return b;
}

void radix_sort_keys(clb_controls_s *ctl, u64 *keys, u64 n, int col_bits, int row_bits) {
    if (n < 2) return;
    const u32 tiles = div_up(n, RS_TILE);
    DevArray<u64> tmp(ctl, n);
    DevArray<u32> counts(ctl, (u64)RS_RADIX * tiles);
    u64 *src = keys, *dst = tmp.ptr;
    std::vector<int> shifts;
    for (int s = 0; s < col_bits; s += 8) shifts.push_back(s);
    for (int s = 0; s < row_bits; s += 8) shifts.push_back(32 + s);
    for (int shift : shifts) {
        k_radix_hist<<<tiles, RS_WG, 0, ctl->stream>>>(src, n, shift, tiles, counts.ptr);
        CLB_LAUNCHED("radix_hist");
        exclusive_scan_u32(ctl, counts.ptr, counts.ptr, (u64)RS_RADIX * tiles, nullptr);
        k_radix_scatter<<<tiles, RS_WG, 0, ctl->stream>>>(src, dst, n, shift, tiles, counts.ptr);
        CLB_LAUNCHED("radix_scatter");
        std::swap(src, dst);
    }
    if (src != keys)
        CLB_HIP(hipMemcpyAsync(keys, src, n * 8, hipMemcpyDeviceToDevice, ctl->stream), COPY_BUFFER_ERROR);
}

void sort_pair_keys(clb_controls_s *ctl, u64 *keys, u64 n, u32 max_row, u32 max_col) {
    radix_sort_keys(ctl, keys, n, bits_for(max_col), bits_for(max_row));
}

__global__ void k_key_max(const u64 *__restrict__ keys, u64 n, u32 *__restrict__ mx) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u32 r = 0, c = 0;
    if (i < n) {
        r = (u32)(keys[i] >> 32);
        c = (u32)keys[i];
    }
    r = dev::wave_max(r);
    c = dev::wave_max(c);
    if (dev::lane() == 0) {
        atomicMax(&mx[0], r);
        atomicMax(&mx[1], c);
    }
}

// flag[i] = key[i] != key[i-1] (first always kept), flag[n] = 0
// Sorted keys -> distinct keys, in tiles of KT_TILE keys.  Key i is kept
// when it differs from key i - 1, and starts a row among the kept keys when
// its row also differs (the previous distinct key is key i - 1).  A count
// pass gives each tile's kept keys and row starts, their scans give each
// tile's output offsets, and the scatter pass places every kept key by its
// rank inside the tile (ballots per wave and item, counts exchanged in LDS).
// Loads are striped (item it of thread t is key it * KT_WG + t of the tile),
// so every load and every run of stores is coalesced.
constexpr int KT_WG = 256, KT_ITEMS = 16;
constexpr u64 KT_TILE = (u64)KT_WG * KT_ITEMS;

// bit 0: kept, bit 1: starts a row among the kept keys
__device__ __forceinline__ u32 key_flags(const u64 *__restrict__ keys, u64 i, u64 k) {
    if (i == 0) return 3u;
    const u64 p = keys[i - 1];
    return k == p ? 0u : ((k >> 32) != (p >> 32) ? 3u : 1u);
}

__global__ __launch_bounds__(KT_WG) void k_keys_count(const u64 *__restrict__ keys, u64 n, u32 *__restrict__ kept,
                                                      u32 *__restrict__ starts) {
    __shared__ u32 red[KT_WG / dev::WAVE + 1];
    const u64 base = (u64)blockIdx.x * KT_TILE;
    u32 cu = 0, cs = 0;
#pragma unroll 4
    for (int it = 0; it < KT_ITEMS; ++it) {
        const u64 i = base + (u64)it * KT_WG + threadIdx.x;
        if (i < n) {
            const u32 f = key_flags(keys, i, keys[i]);
            cu += f & 1u;
            cs += f >> 1;
        }
    }
    const u32 tu = dev::block_sum<u32, KT_WG>(cu, red);
    const u32 ts = dev::block_sum<u32, KT_WG>(cs, red);
    if (threadIdx.x == 0) {
        kept[blockIdx.x] = tu;
        starts[blockIdx.x] = ts;
    }
}

// DCSR: cols[p] = column of the p-th kept key; for the q-th row start,
// rows[q] = its row and rpt[q] = p; rpt[nzr] = m.  Pairs (!DCSR): rows[p] =
// row of the p-th kept key.
template <bool DCSR>
__global__ __launch_bounds__(KT_WG) void k_keys_scatter(const u64 *__restrict__ keys, u64 n,
                                                        const u32 *__restrict__ koff, const u32 *__restrict__ soff,
                                                        const u32 *__restrict__ totals, u32 *__restrict__ cols,
                                                        u32 *__restrict__ rows, u32 *__restrict__ rpt) {
    constexpr int NW = KT_WG / dev::WAVE;
    __shared__ u32 cu[KT_ITEMS][NW], cs[KT_ITEMS][NW];
    const u64 base = (u64)blockIdx.x * KT_TILE;
    const u64 *kt = keys + base;  // the tile, indexed by 32-bit offsets
    const u32 nt = n - base < KT_TILE ? (u32)(n - base) : (u32)KT_TILE;
    const u32 w = dev::wave_id();
    const u32 l = (u32)dev::lane();
    u32 fl = 0;  // bit it: kept, bit 16 + it: row start
    // pass 1: flags; per wave and item the counts go to LDS
#pragma unroll 4
    for (int it = 0; it < KT_ITEMS; ++it) {
        const u32 o = (u32)it * KT_WG + threadIdx.x;
        const u32 f = o < nt ? key_flags(keys, base + o, kt[o]) : 0u;
        fl |= (f & 1u) << it;
        fl |= (f >> 1) << (16 + it);
        const u64 bu = __ballot(f & 1u), bs = __ballot(f >> 1);
        if (l == 0) {
            cu[it][w] = (u32)__popcll(bu);
            cs[it][w] = (u32)__popcll(bs);
        }
    }
    __syncthreads();
    // pass 2: ranks; the kept keys are read again (from cache)
    u32 au = koff[blockIdx.x], as = soff[blockIdx.x];
#pragma unroll 4
    for (int it = 0; it < KT_ITEMS; ++it) {
        u32 lu = 0, ls = 0, tu = 0, ts = 0;
#pragma unroll
        for (int v = 0; v < NW; ++v) {
            const u32 x = cu[it][v], y = cs[it][v];
            tu += x;
            ts += y;
            lu += (u32)v < w ? x : 0u;
            ls += (u32)v < w ? y : 0u;
        }
        const bool kept = (fl >> it) & 1u, start = (fl >> (16 + it)) & 1u;
        const u64 bu = __ballot(kept), bs = __ballot(start);
        if (kept) {
            const u64 k = kt[(u32)it * KT_WG + threadIdx.x];
            const u32 p = au + lu + dev::lanes_below(bu);
            cols[p] = (u32)k;
            if (!DCSR) rows[p] = (u32)(k >> 32);
            if (DCSR && start) {
                const u32 q = as + ls + dev::lanes_below(bs);
                rows[q] = (u32)(k >> 32);
                rpt[q] = p;
            }
        }
        au += tu;
        as += ts;
    }
    if (DCSR && blockIdx.x == 0 && threadIdx.x == 0) rpt[totals[1]] = totals[0];
}

// Counts and offsets of the kept keys and row starts: tot[0] = kept keys,
// tot[1] = row starts.
static void keys_offsets(clb_controls_s *ctl, const u64 *keys, u64 n, DevArray<u32> &koff, DevArray<u32> &soff,
                         DevArray<u32> &tot, u32 &m, u32 &nzr) {
    if (n >= 0xFFFFFFFFull) raise(INVALID_VALUE, "more keys than uint32 indices allow");
    const u64 tiles = (n + KT_TILE - 1) / KT_TILE;
    koff.reset(ctl, tiles);
    soff.reset(ctl, tiles);
    tot.reset(ctl, 2);
    k_keys_count<<<(u32)tiles, KT_WG, 0, ctl->stream>>>(keys, n, koff.ptr, soff.ptr);
    CLB_LAUNCHED("keys_count");
    exclusive_scan_u32(ctl, koff.ptr, koff.ptr, tiles, tot.ptr);
    exclusive_scan_u32(ctl, soff.ptr, soff.ptr, tiles, tot.ptr + 1);
    u32 h[2];
    readback(ctl, h, tot.ptr, 8);
    m = h[0];
    nzr = h[1];
}

void keys_to_dcsr(clb_controls_s *ctl, const u64 *keys, u64 n, clb_dcsr &out) {
    clb_dcsr r{};
    r.nrows = out.nrows;
    r.ncols = out.ncols;
    if (n == 0) {
        out = r;
        return;
    }
    DevArray<u32> koff, soff, tot;
    u32 m = 0, nzr = 0;
    keys_offsets(ctl, keys, n, koff, soff, tot, m, nzr);
    try {
        r.cols = buffer_new(ctl, 4ull * m);
        r.rows = buffer_new(ctl, 4ull * nzr);
        r.rpt = buffer_new(ctl, 4ull * (nzr + 1ull));
        k_keys_scatter<true><<<(u32)koff.n, KT_WG, 0, ctl->stream>>>(keys, n, koff.ptr, soff.ptr, tot.ptr,
                                                                      u32p(r.cols), u32p(r.rows), u32p(r.rpt));
        CLB_LAUNCHED("keys_scatter");
        sync(ctl);
    } catch (...) {
        clb_dcsr_release(&r);
        throw;
    }
    r.nnz = m;
    r.nzr = nzr;
    out = r;
}

void pack_pairs(clb_controls_s *ctl, const u32 *rows, const u32 *cols, u64 n, u64 *keys) {
    if (!n) return;
    k_pack_pairs<<<div_up(n, 256), 256, 0, ctl->stream>>>(rows, cols, n, keys);
    CLB_LAUNCHED("pack_pairs");
}

void unpack_pairs(clb_controls_s *ctl, const u64 *keys, u64 n, u32 *rows, u32 *cols) {
    if (!n) return;
    k_unpack_pairs<<<div_up(n, 256), 256, 0, ctl->stream>>>(keys, n, rows, cols);
    CLB_LAUNCHED("unpack_pairs");
}

// Sorted keys -> deduplicated (rows, cols) buffers.
u32 unique_keys_to_pairs(clb_controls_s *ctl, const u64 *keys, u64 n, clb_buffer_s **rows_out,
                         clb_buffer_s **cols_out) {
    *rows_out = *cols_out = nullptr;
    if (n == 0) {
        *rows_out = buffer_new(ctl, 4);
        *cols_out = buffer_new(ctl, 4);
        return 0;
    }
    DevArray<u32> koff, soff, tot;
    u32 m = 0, nzr = 0;
    keys_offsets(ctl, keys, n, koff, soff, tot, m, nzr);
    *rows_out = buffer_new(ctl, 4ull * (m ? m : 1));
    try {
        *cols_out = buffer_new(ctl, 4ull * (m ? m : 1));
    } catch (...) {
        buffer_release(*rows_out);
        *rows_out = nullptr;
        throw;
    }
    k_keys_scatter<false><<<(u32)koff.n, KT_WG, 0, ctl->stream>>>(keys, n, koff.ptr, soff.ptr, tot.ptr,
                                                                   u32p(*cols_out), u32p(*rows_out), nullptr);
    CLB_LAUNCHED("keys_scatter");
    return m;
}

void key_bounds(clb_controls_s *ctl, const u64 *keys, u64 n, u32 *max_row, u32 *max_col) {
    DevArray<u32> mx(ctl, 2);
    fill_u32(ctl, mx.ptr, 2, 0);
    if (n) {
        k_key_max<<<div_up(n, 256), 256, 0, ctl->stream>>>(keys, n, mx.ptr);
        CLB_LAUNCHED("key_max");
    }
    u32 h[2];
    readback(ctl, h, mx.ptr, 8);
    *max_row = h[0];
    *max_col = h[1];
}

// COO construction: matrix_coo.cpp:17-47 (sort if !sorted, dedup if
// !noDuplicates).  rows/cols are consumed (their references move to out).
void coo_finish(clb_controls_s *ctl, clb_buffer_s *rows, clb_buffer_s *cols, u32 nrows, u32 ncols, u32 nnz,
                bool sorted, bool nodup, clb_coo &out) {
    out = clb_coo{};
    out.nrows = nrows;
    out.ncols = ncols;
    out.rows = rows;
    out.cols = cols;
    out.nnz = nnz;
    if (nnz == 0 || (sorted && nodup)) return;
    DevArray<u64> keys(ctl, nnz);
    pack_pairs(ctl, u32p(rows), u32p(cols), nnz, keys.ptr);
    if (!sorted) {
        u32 mr, mc;
        key_bounds(ctl, keys.ptr, nnz, &mr, &mc);
        sort_pair_keys(ctl, keys.ptr, nnz, mr, mc);
    }
    if (!nodup) {
        clb_buffer_s *nr = nullptr, *nc = nullptr;
        const u32 m = unique_keys_to_pairs(ctl, keys.ptr, nnz, &nr, &nc);
        sync(ctl);
        buffer_release(rows);
        buffer_release(cols);
        out.rows = nr;
        out.cols = nc;
        out.nnz = m;
    } else {
        unpack_pairs(ctl, keys.ptr, nnz, u32p(rows), u32p(cols));
        sync(ctl);
    }
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_coo_from_host(clb_controls ctl, const uint32_t *rows, const uint32_t *cols, uint32_t nrows,
                             uint32_t ncols, uint32_t nnz, int sorted, int no_duplicates, clb_coo *out) {
    return run_guarded([&] {
        clb_coo r{};
        r.nrows = nrows;
        r.ncols = ncols;
        if (nnz == 0) {  // matrix_coo.cpp:29
            *out = r;
            return;
        }
        clb_buffer_s *rb = buffer_new(ctl, 4ull * nnz);
        clb_buffer_s *cbuf = nullptr;
        try {
            cbuf = buffer_new(ctl, 4ull * nnz);
            CLB_HIP(hipMemcpyAsync(rb->ptr, rows, 4ull * nnz, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
            CLB_HIP(hipMemcpyAsync(cbuf->ptr, cols, 4ull * nnz, hipMemcpyHostToDevice, ctl->stream),
                    WRITE_BUFFER_ERROR);
            sync(ctl);
        } catch (...) {
            buffer_release(rb);
            buffer_release(cbuf);
            throw;
        }
        coo_finish(ctl, rb, cbuf, nrows, ncols, nnz, sorted != 0, no_duplicates != 0, r);
        *out = r;
    });
}

clb_status clb_coo_from_buffers(clb_controls ctl, uint32_t nrows, uint32_t ncols, uint32_t nnz, clb_buffer rows,
                                clb_buffer cols, int sorted, int no_duplicates, clb_coo *out) {
    return run_guarded([&] {
        clb_coo r{};
        if (nnz && (!rows || !cols || rows->bytes < 4ull * nnz || cols->bytes < 4ull * nnz))
            raise(INVALID_ARGUMENT, "matrix_coo: buffers smaller than nnz");
        // matrix_coo.cpp:59: the matrix shares the given buffers
        buffer_retain(rows);
        buffer_retain(cols);
        if (!sorted && nnz) {  // sorting happens in place, as coo::sort_arrays does
            DevArray<u64> keys(ctl, nnz);
            pack_pairs(ctl, u32p(rows), u32p(cols), nnz, keys.ptr);
            u32 mr, mc;
            key_bounds(ctl, keys.ptr, nnz, &mr, &mc);
            sort_pair_keys(ctl, keys.ptr, nnz, mr, mc);
            unpack_pairs(ctl, keys.ptr, nnz, u32p(rows), u32p(cols));
            sync(ctl);
        }
        coo_finish(ctl, nnz ? rows : nullptr, nnz ? cols : nullptr, nrows, ncols, nnz, true, no_duplicates != 0, r);
        if (!nnz) {
            buffer_release(rows);
            buffer_release(cols);
        }
        *out = r;
    });
}

clb_status clb_coo_sort_arrays(clb_controls ctl, clb_buffer rows, clb_buffer cols, uint32_t n) {
    return run_guarded([&] {
        if (n == 0) return;
        if (!rows || !cols || rows->bytes < 4ull * n || cols->bytes < 4ull * n)
            raise(INVALID_ARGUMENT, "sort_arrays: buffers smaller than n");
        DevArray<u64> keys(ctl, n);
        pack_pairs(ctl, u32p(rows), u32p(cols), n, keys.ptr);
        u32 mr, mc;
        key_bounds(ctl, keys.ptr, n, &mr, &mc);
        sort_pair_keys(ctl, keys.ptr, n, mr, mc);
        unpack_pairs(ctl, keys.ptr, n, u32p(rows), u32p(cols));
        sync(ctl);
    });
}

}  // extern "C"
