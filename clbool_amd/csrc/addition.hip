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

constexpr int MERGE_ITEMS = 8;

__global__ __launch_bounds__(256) void k_merge_keys(const u64 *__restrict__ A, u64 na, const u64 *__restrict__ B,
                                                    u64 nb, u64 *__restrict__ out) {
    const u64 total = na + nb;
    const u64 d0 = ((u64)blockIdx.x * 256 + threadIdx.x) * MERGE_ITEMS;
    if (d0 >= total) return;
    // merge_path.cl:38-90: elements of A among the first d0 outputs (A first on ties)
    u64 lo = d0 > nb ? d0 - nb : 0, hi = d0 < na ? d0 : na;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (A[mid] <= B[d0 - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    u64 i = lo, j = d0 - lo;
#pragma unroll
    for (int k = 0; k < MERGE_ITEMS; ++k) {
        if (d0 + k >= total) break;
        const bool take_a = i < na && (j >= nb || A[i] <= B[j]);
        out[d0 + k] = take_a ? A[i++] : B[j++];
    }
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

// Merge two sorted key lists and drop duplicates -> (rows, cols) buffers.
static u32 union_keys(clb_controls_s *ctl, const u64 *ka, u64 na, const u64 *kb, u64 nb, clb_buffer_s **rows,
                      clb_buffer_s **cols) {
    DevArray<u64> merged(ctl, na + nb);
    const u64 threads = (na + nb + MERGE_ITEMS - 1) / MERGE_ITEMS;
    k_merge_keys<<<div_up(threads, 256), 256, 0, ctl->stream>>>(ka, na, kb, nb, merged.ptr);
    CLB_LAUNCHED("merge");
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
        clb_buffer_s *rows = nullptr;
        const u32 m = union_keys(ctl, ka.ptr, a->nnz, kb.ptr, b->nnz, &rows, &r.cols);
        try {
            rows_to_dcsr(ctl, u32p(rows), m, &r.rpt, &r.rows, &r.nzr);
        } catch (...) {
            buffer_release(rows);
            clb_dcsr_release(&r);
            throw;
        }
        buffer_release(rows);
        r.nnz = m;
        *out = r;
    });
}

}  // extern "C"
