// addition.hip -- boolean matrix addition (union) by merge path.
//
// coo::matrix_addition (src/coo/coo_matrix_addition.cpp:32-73) merges two
// sorted (row, col) lists with merge_path.cl:182-261 and then drops
// duplicates (reduce_duplicates2, matrix_coo.cpp:117-170).  Here the pairs
// are 64-bit keys row << 32 | col; each thread finds its start on the merge
// path by one binary search over the cross diagonal and merges 8 outputs
// sequentially; duplicates are then removed by flag, scan, scatter.
//
// csr::matrix_addition (src/csr/csr_addition.cpp:26-168) bins rows by length
// and merges each row in a workgroup.  Because CSR rows with sorted columns
// are already one globally sorted key list, the same flat merge handles it
// with no per-row load imbalance; the CSR row pointer is rebuilt by one
// lower-bound search per row.  The DCSR union (the X += T step of a closure)
// reuses the same engine.
#include "sort.hpp"

#include "device_utils.hpp"

namespace clb {

// Merge of two ascending key lists, in tiles of MG_TILE outputs.  A first
// pass finds where every tile boundary cuts A (merge_path.cl:38-90: the
// number of A elements among the first d outputs, A first on ties), one
// thread per boundary.  Each workgroup then loads its tile's pieces of A and B
// into LDS with coalesced loads, every thread merges MG_ITEMS consecutive
// outputs from LDS (its own merge-path search within the tile), and the tile
// is written back with coalesced stores.
constexpr int MG_WG = 256, MG_ITEMS = 8, MG_TILE = MG_WG * MG_ITEMS;

__global__ void k_merge_partition(const u64 *__restrict__ A, u64 na, const u64 *__restrict__ B, u64 nb, u64 ntiles,
                                  u64 *__restrict__ split) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    const u64 total = na + nb;
    const u64 d = t * MG_TILE < total ? t * MG_TILE : total;
    u64 lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    split[t] = lo;
}

__global__ __launch_bounds__(MG_WG) void k_merge_tiles(const u64 *__restrict__ A, u64 na, const u64 *__restrict__ B,
                                                       u64 nb, const u64 *__restrict__ split,
                                                       u64 *__restrict__ out) {
    __shared__ u64 s[MG_TILE];
    const u64 total = na + nb;
    const u64 d0 = (u64)blockIdx.x * MG_TILE;
    const u64 d1 = d0 + MG_TILE < total ? d0 + MG_TILE : total;
    const u64 a0 = split[blockIdx.x], a1 = split[blockIdx.x + 1];
    const u64 b0 = d0 - a0;
    const u32 la = (u32)(a1 - a0), n = (u32)(d1 - d0), lb = n - la;
    for (u32 t = threadIdx.x; t < n; t += MG_WG) s[t] = t < la ? A[a0 + t] : B[b0 + (t - la)];
    __syncthreads();
    const u32 k0 = threadIdx.x * MG_ITEMS;
    u64 v[MG_ITEMS];
    if (k0 < n) {
        u32 lo = k0 > lb ? k0 - lb : 0, hi = k0 < la ? k0 : la;
        while (lo < hi) {
            const u32 mid = (lo + hi) >> 1;
            if (s[mid] <= s[la + k0 - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        u32 i = lo, j = k0 - lo;
#pragma unroll
        for (int k = 0; k < MG_ITEMS; ++k) {
            const bool take_a = i < la && (j >= lb || s[i] <= s[la + j]);
            v[k] = s[take_a ? i : (j < lb ? la + j : n - 1)];  // n - 1: past the tile, never stored
            i += take_a;
            j += !take_a;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MG_ITEMS; ++k)
        if (k0 + k < n) s[k0 + k] = v[k];
    __syncthreads();
    for (u32 t = threadIdx.x; t < n; t += MG_WG) out[d0 + t] = s[t];
}

// Keys of a CSR matrix, one wave per row.
__global__ __launch_bounds__(256) void k_csr_keys(const u32 *__restrict__ rpt, const u32 *__restrict__ cols,
                                                  u32 nrows, u32 limit, u64 *__restrict__ keys) {
    const u32 r = blockIdx.x * 4 + dev::wave_id();
    if (r >= nrows) return;
    const u32 e = rpt[r + 1] < limit ? rpt[r + 1] : limit;
    for (u32 j = rpt[r] + dev::lane(); j < e; j += 64) keys[j] = ((u64)r << 32) | cols[j];
}

// Keys of a DCSR matrix, one wave per stored row.
__global__ __launch_bounds__(256) void k_dcsr_keys(const u32 *__restrict__ rpt, const u32 *__restrict__ rows,
                                                   const u32 *__restrict__ cols, u32 nzr, u64 *__restrict__ keys) {
    const u32 r = blockIdx.x * 4 + dev::wave_id();
    if (r >= nzr) return;
    const u64 hi = (u64)rows[r] << 32;
    const u32 e = rpt[r + 1];
    for (u32 j = rpt[r] + dev::lane(); j < e; j += 64) keys[j] = hi | cols[j];
}

// CSR row pointer of sorted rows: rpt[r] = lower_bound(rows, r), r <= nrows.
__global__ void k_rows_lower_bound(const u32 *__restrict__ rows, u32 m, u32 nrows, u32 *__restrict__ rpt) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    u32 lo = 0, hi = m;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo) / 2;
        if (rows[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    rpt[r] = lo;
}

void merge_keys(clb_controls_s *ctl, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out) {
    if (na + nb == 0) return;
    const u64 tiles = (na + nb + MG_TILE - 1) / MG_TILE;
    if (tiles >= 0xFFFFFFFFull) raise(INVALID_VALUE, "merge larger than the grid allows");
    DevArray<u64> split(ctl, tiles + 1);
    k_merge_partition<<<div_up(tiles + 1, 256), 256, 0, ctl->stream>>>(a, na, b, nb, tiles, split.ptr);
    CLB_LAUNCHED("merge_partition");
    k_merge_tiles<<<(u32)tiles, MG_WG, 0, ctl->stream>>>(a, na, b, nb, split.ptr, out);
    CLB_LAUNCHED("merge");
}

// Merge two sorted key lists and drop duplicates -> (rows, cols) buffers.
static u32 union_keys(clb_controls_s *ctl, const u64 *ka, u64 na, const u64 *kb, u64 nb, clb_buffer_s **rows,
                      clb_buffer_s **cols) {
    DevArray<u64> merged(ctl, na + nb);
    merge_keys(ctl, ka, na, kb, nb, merged.ptr);
    return unique_keys_to_pairs(ctl, merged.ptr, na + nb, rows, cols);
}

static void check_same_shape(u32 ar, u32 ac, u32 br, u32 bc) {
    // coo_matrix_addition.cpp:36-41, csr_addition.cpp:28-33
    if (ar != br || ac != bc) {
        std::string s = "Invalid matrixes size! a: " + std::to_string(ar) + " x " + std::to_string(ac) +
                        ", b: " + std::to_string(br) + " x " + std::to_string(bc);
        raise(INVALID_ARGUMENT, s);
    }
}

static clb_buffer_s *copy_buffer(clb_controls_s *ctl, const clb_buffer_s *src, u64 bytes) {
    clb_buffer_s *b = buffer_new(ctl, bytes);
    hipError_t e = hipMemcpyAsync(b->ptr, src->ptr, bytes, hipMemcpyDeviceToDevice, ctl->stream);
    if (e != hipSuccess) {
        buffer_release(b);
        check_hip(e, COPY_BUFFER_ERROR, "copy buffer");
    }
    return b;
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_coo_addition(clb_controls ctl, const clb_coo *a, const clb_coo *b, clb_coo *out) {
    return run_guarded([&] {
        check_same_shape(a->nrows, a->ncols, b->nrows, b->ncols);
        clb_coo r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz == 0 && b->nnz == 0) {  // :43-46
            *out = r;
            return;
        }
        if (a->nnz == 0 || b->nnz == 0) {  // :48-64: a copy of the filled operand
            const clb_coo *f = a->nnz ? a : b;
            r.rows = copy_buffer(ctl, f->rows, 4ull * f->nnz);
            try {
                r.cols = copy_buffer(ctl, f->cols, 4ull * f->nnz);
                sync(ctl);
            } catch (...) {
                clb_coo_release(&r);
                throw;
            }
            r.nnz = f->nnz;
            *out = r;
            return;
        }
        DevArray<u64> ka(ctl, a->nnz), kb(ctl, b->nnz);
        pack_pairs(ctl, u32p(a->rows), u32p(a->cols), a->nnz, ka.ptr);
        pack_pairs(ctl, u32p(b->rows), u32p(b->cols), b->nnz, kb.ptr);
        r.nnz = union_keys(ctl, ka.ptr, a->nnz, kb.ptr, b->nnz, &r.rows, &r.cols);
        sync(ctl);
        *out = r;
    });
}

clb_status clb_csr_addition(clb_controls ctl, const clb_csr *a, const clb_csr *b, clb_csr *out) {
    return run_guarded([&] {
        check_same_shape(a->nrows, a->ncols, b->nrows, b->ncols);
        clb_csr r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz == 0 && b->nnz == 0) {  // :35-38
            *out = r;
            return;
        }
        if (a->nnz == 0 || b->nnz == 0) {  // :40-55
            const clb_csr *f = a->nnz ? a : b;
            r.rpt = copy_buffer(ctl, f->rpt, 4ull * (f->nrows + 1ull));
            try {
                r.cols = copy_buffer(ctl, f->cols, 4ull * f->nnz);
                sync(ctl);
            } catch (...) {
                clb_csr_release(&r);
                throw;
            }
            r.nnz = f->nnz;
            *out = r;
            return;
        }
        // the rows [0, nrows) hold rpt[nrows] entries; like the reference's
        // per-row kernels (csr_addition.cl), only those are merged
        u32 na = 0, nb = 0;
        readback(ctl, &na, u32p(a->rpt) + a->nrows, 4);
        readback(ctl, &nb, u32p(b->rpt) + b->nrows, 4);
        na = na < a->nnz ? na : a->nnz;
        nb = nb < b->nnz ? nb : b->nnz;
        DevArray<u64> ka(ctl, na ? na : 1), kb(ctl, nb ? nb : 1);
        k_csr_keys<<<div_up(a->nrows, 4), 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(a->cols), a->nrows, na, ka.ptr);
        k_csr_keys<<<div_up(b->nrows, 4), 256, 0, ctl->stream>>>(u32p(b->rpt), u32p(b->cols), b->nrows, nb, kb.ptr);
        CLB_LAUNCHED("csr_keys");
        clb_buffer_s *rows = nullptr;
        const u32 m = union_keys(ctl, ka.ptr, na, kb.ptr, nb, &rows, &r.cols);
        try {
            r.rpt = buffer_new(ctl, 4ull * (a->nrows + 1ull));
            k_rows_lower_bound<<<div_up(a->nrows + 1ull, 256), 256, 0, ctl->stream>>>(u32p(rows), m, a->nrows,
                                                                                     u32p(r.rpt));
            CLB_LAUNCHED("rows_lower_bound");
            sync(ctl);
        } catch (...) {
            buffer_release(rows);
            clb_csr_release(&r);
            throw;
        }
        buffer_release(rows);
        r.nnz = m;
        *out = r;
    });
}

clb_status clb_dcsr_addition(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out) {
    return run_guarded([&] {
        check_same_shape(a->nrows, a->ncols, b->nrows, b->ncols);
        clb_dcsr r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz == 0 && b->nnz == 0) {
            *out = r;
            return;
        }
        if (a->nnz == 0 || b->nnz == 0) {
            const clb_dcsr *f = a->nnz ? a : b;
            try {
                r.rpt = copy_buffer(ctl, f->rpt, 4ull * (f->nzr + 1ull));
                r.rows = copy_buffer(ctl, f->rows, 4ull * f->nzr);
                r.cols = copy_buffer(ctl, f->cols, 4ull * f->nnz);
                sync(ctl);
            } catch (...) {
                clb_dcsr_release(&r);
                throw;
            }
            r.nnz = f->nnz;
            r.nzr = f->nzr;
            *out = r;
            return;
        }
        DevArray<u64> ka(ctl, a->nnz), kb(ctl, b->nnz);
        k_dcsr_keys<<<div_up(a->nzr, 4), 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(a->rows), u32p(a->cols), a->nzr,
                                                                 ka.ptr);
        k_dcsr_keys<<<div_up(b->nzr, 4), 256, 0, ctl->stream>>>(u32p(b->rpt), u32p(b->rows), u32p(b->cols), b->nzr,
                                                                 kb.ptr);
        CLB_LAUNCHED("dcsr_keys");
        DevArray<u64> merged(ctl, (u64)a->nnz + b->nnz);
        merge_keys(ctl, ka.ptr, a->nnz, kb.ptr, b->nnz, merged.ptr);
        keys_to_dcsr(ctl, merged.ptr, (u64)a->nnz + b->nnz, r);
        *out = r;
    });
}

}  // extern "C"
