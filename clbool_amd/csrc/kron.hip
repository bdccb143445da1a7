// kron.hip -- DCSR Kronecker product C = A (x) B.
//
// dcsr::kronecker_product (src/dcsr/dcsr_kronecker_product.cpp:6-65) with
// count_nnz_per_row / calculate_kronecker_product (dcsr_kronecker.cl:38-98):
// output row r = ia * b_nzr + ib has |A_ia| * |B_ib| entries, row id
// a_row * b_nrows + b_row and columns a_col * b_ncols + b_col (already in
// ascending order).  The reference finds each output's row by a binary search
// per nonzero; here a thread searches once and then walks 8 consecutive
// outputs, so the kernel is a streaming write.  Sizes are checked in 64 bits:
// the reference computes nnz, nzr and the shape as wrapping uint32 products
// (:10-13), which this library reports as an error instead.
#include "clb_internal.hpp"

#include "device_utils.hpp"

namespace clb {

constexpr int KRON_ITEMS = 8;

__global__ void k_kron_rows(const u32 *__restrict__ a_rpt, const u32 *__restrict__ b_rpt,
                            const u32 *__restrict__ a_rows, const u32 *__restrict__ b_rows, u32 c_nzr, u32 b_nzr,
                            u32 b_nrows, u32 *__restrict__ c_rpt, u32 *__restrict__ c_rows) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) c_rpt[c_nzr] = 0;
    if (r >= c_nzr) return;
    const u32 ia = r / b_nzr, ib = r % b_nzr;
    c_rpt[r] = (a_rpt[ia + 1] - a_rpt[ia]) * (b_rpt[ib + 1] - b_rpt[ib]);
    c_rows[r] = a_rows[ia] * b_nrows + b_rows[ib];
}

__global__ __launch_bounds__(256) void k_kron_fill(const u32 *__restrict__ c_rpt, u32 c_nzr, u32 c_nnz,
                                                   const u32 *__restrict__ a_rpt, const u32 *__restrict__ b_rpt,
                                                   const u32 *__restrict__ a_cols, const u32 *__restrict__ b_cols,
                                                   u32 b_nzr, u32 b_ncols, u32 *__restrict__ c_cols) {
    const u64 e0 = ((u64)blockIdx.x * 256 + threadIdx.x) * KRON_ITEMS;
    if (e0 >= c_nnz) return;
    // row of e0: last r with c_rpt[r] <= e0 (rows with no entries are skipped)
    u32 lo = 0, hi = c_nzr;
    while (hi - lo > 1) {
        const u32 mid = lo + (hi - lo) / 2;
        if (c_rpt[mid] <= e0) lo = mid;
        else hi = mid;
    }
    u32 r = lo;
    while (c_rpt[r + 1] <= e0) ++r;
    u32 ia = r / b_nzr, ib = r % b_nzr;
    u32 blen = b_rpt[ib + 1] - b_rpt[ib];
    u32 off = (u32)e0 - c_rpt[r];
    u32 ka = off / blen, kb = off % blen;
    for (int k = 0; k < KRON_ITEMS; ++k) {
        const u64 e = e0 + k;
        if (e >= c_nnz) break;
        c_cols[e] = a_cols[a_rpt[ia] + ka] * b_ncols + b_cols[b_rpt[ib] + kb];
        if (++kb == blen) {
            kb = 0;
            if (++ka == a_rpt[ia + 1] - a_rpt[ia] && e + 1 < c_nnz) {
                ka = 0;
                do {  // next row that has entries
                    ++r;
                } while (c_rpt[r + 1] == c_rpt[r]);
                ia = r / b_nzr;
                ib = r % b_nzr;
                blen = b_rpt[ib + 1] - b_rpt[ib];
            }
        }
    }
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_dcsr_kronecker(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out) {
    return run_guarded([&] {
        const u64 c_nnz = (u64)a->nnz * b->nnz;
        const u64 c_nzr = (u64)a->nzr * b->nzr;
        const u64 c_nrows = (u64)a->nrows * b->nrows;
        const u64 c_ncols = (u64)a->ncols * b->ncols;
        if (c_nrows > 0xFFFFFFFFull || c_ncols > 0xFFFFFFFFull)
            raise(INVALID_VALUE, "kronecker product shape exceeds uint32 indices");
        clb_dcsr r{};
        r.nrows = (u32)c_nrows;
        r.ncols = (u32)c_ncols;
        if (c_nnz == 0) {  // :15-18
            *out = r;
            return;
        }
        if (c_nnz > 0xFFFFFFFFull) raise(INVALID_VALUE, "kronecker product has more than 2^32-1 nonzeros");
        try {
            r.rpt = buffer_new(ctl, 4ull * (c_nzr + 1));
            r.rows = buffer_new(ctl, 4ull * c_nzr);
            r.cols = buffer_new(ctl, 4ull * c_nnz);
            k_kron_rows<<<div_up(c_nzr + 1, 256), 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(b->rpt), u32p(a->rows),
                                                                         u32p(b->rows), (u32)c_nzr, b->nzr, b->nrows,
                                                                         u32p(r.rpt), u32p(r.rows));
            CLB_LAUNCHED("count_nnz_per_row");
            DevArray<u32> tot(ctl, 1);
            exclusive_scan_u32(ctl, u32p(r.rpt), u32p(r.rpt), c_nzr + 1, tot.ptr);
            const u64 threads = (c_nnz + KRON_ITEMS - 1) / KRON_ITEMS;
            k_kron_fill<<<div_up(threads, 256), 256, 0, ctl->stream>>>(u32p(r.rpt), (u32)c_nzr, (u32)c_nnz,
                                                                       u32p(a->rpt), u32p(b->rpt), u32p(a->cols),
                                                                       u32p(b->cols), b->nzr, b->ncols, u32p(r.cols));
            CLB_LAUNCHED("calculate_kronecker_product");
            u32 total = 0;
            readback(ctl, &total, tot.ptr, 4);
            if (total != c_nnz) raise(RUN_KERNEL_FAILURE, "kronecker row sizes do not add up to nnz");  // :46
        } catch (...) {
            (void)hipStreamSynchronize(ctl->stream);
            clb_dcsr_release(&r);
            throw;
        }
        r.nnz = (u32)c_nnz;
        r.nzr = (u32)c_nzr;
        *out = r;
    });
}

}  // extern "C"
