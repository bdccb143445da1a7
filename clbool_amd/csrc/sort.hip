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
__global__ void k_unique_flags(const u64 *__restrict__ keys, u64 n, u32 *__restrict__ flags) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
    if (i == n) flags[n] = 0;
}

__global__ void k_unique_scatter(const u64 *__restrict__ keys, u64 n, const u32 *__restrict__ pos,
                                 u32 *__restrict__ rows, u32 *__restrict__ cols) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pos[i] != pos[i + 1]) {
        rows[pos[i]] = (u32)(keys[i] >> 32);
        cols[pos[i]] = (u32)keys[i];
    }
}

__global__ void k_unique_key_scatter(const u64 *__restrict__ keys, u64 n, const u32 *__restrict__ pos,
                                     u64 *__restrict__ out) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pos[i] != pos[i + 1]) out[pos[i]] = keys[i];
}

u32 unique_keys(clb_controls_s *ctl, const u64 *keys, u64 n, DevArray<u64> &out) {
    if (n == 0) {
        out.reset(ctl, 1);
        return 0;
    }
    DevArray<u32> pos(ctl, n + 1);
    k_unique_flags<<<div_up(n + 1, 256), 256, 0, ctl->stream>>>(keys, n, pos.ptr);
    CLB_LAUNCHED("unique_flags");
    DevArray<u32> tot(ctl, 1);
    exclusive_scan_u32(ctl, pos.ptr, pos.ptr, n + 1, tot.ptr);
    u32 m = 0;
    readback(ctl, &m, tot.ptr, 4);
    out.reset(ctl, m ? m : 1);
    k_unique_key_scatter<<<div_up(n, 256), 256, 0, ctl->stream>>>(keys, n, pos.ptr, out.ptr);
    CLB_LAUNCHED("unique_key_scatter");
    return m;
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
    DevArray<u32> pos(ctl, n + 1);
    k_unique_flags<<<div_up(n + 1, 256), 256, 0, ctl->stream>>>(keys, n, pos.ptr);
    CLB_LAUNCHED("unique_flags");
    DevArray<u32> tot(ctl, 1);
    exclusive_scan_u32(ctl, pos.ptr, pos.ptr, n + 1, tot.ptr);
    u32 m = 0;
    readback(ctl, &m, tot.ptr, 4);
    *rows_out = buffer_new(ctl, 4ull * (m ? m : 1));
    *cols_out = buffer_new(ctl, 4ull * (m ? m : 1));
    k_unique_scatter<<<div_up(n, 256), 256, 0, ctl->stream>>>(keys, n, pos.ptr, u32p(*rows_out), u32p(*cols_out));
    CLB_LAUNCHED("unique_scatter");
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
