// kron.hip -- DCSR Kronecker product C = A (x) B.
//
// dcsr::kronecker_product (src/dcsr/dcsr_kronecker_product.cpp:6-65) with
// count_nnz_per_row / calculate_kronecker_product (dcsr_kronecker.cl:38-98):
// output row r = ia * b_nzr + ib has |A_ia| * |B_ib| entries, row id
// a_row * b_nrows + b_row and columns a_col * b_ncols + b_col (already in
// ascending order).  Output rows are A-row-major, so row (ia, ib) starts at
// a_rpt[ia] * b_nnz + |A_ia| * b_rpt[ib]: the row pointer is a closed form,
// no scan.  The reference finds each output's row by a binary search per
// nonzero; here the work is cut into tiles of B nonzeros (one workgroup per
// tile and A row): a partition pass finds each tile's first B row, the tile's
// slice of b_rpt is staged in LDS, and each thread places its B nonzero once
// per column of the A row -- consecutive threads write consecutive columns.
// Sizes are checked in 64 bits: the reference computes nnz, nzr and the
// shape as wrapping uint32 products (:10-13), which this library reports as
// an error instead.
#include "clb_internal.hpp"

#include "device_utils.hpp"

namespace clb {

constexpr u32 KRON_WG = 256, KRON_ITEMS = 4, KRON_TILE = KRON_WG * KRON_ITEMS;

__global__ void k_kron_rows(const u32 *__restrict__ a_rpt, const u32 *__restrict__ b_rpt,
                            const u32 *__restrict__ a_rows, const u32 *__restrict__ b_rows, u32 c_nzr, u32 b_nzr,
                            u32 b_nnz, u32 b_nrows, u32 c_nnz, u32 *__restrict__ c_rpt, u32 *__restrict__ c_rows) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) c_rpt[c_nzr] = c_nnz;
    if (r >= c_nzr) return;
    const u32 ia = r / b_nzr, ib = r % b_nzr;
    c_rpt[r] = a_rpt[ia] * b_nnz + (a_rpt[ia + 1] - a_rpt[ia]) * b_rpt[ib];
    c_rows[r] = a_rows[ia] * b_nrows + b_rows[ib];
}

// first[t] = the B row holding B nonzero t * KRON_TILE.
__global__ void k_kron_tiles(const u32 *__restrict__ b_rpt, u32 b_nzr, u32 ntiles, u32 *__restrict__ first) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const u32 e = t * KRON_TILE;
    u32 lo = 0, hi = b_nzr;  // last row with b_rpt[row] <= e
    while (hi - lo > 1) {
        const u32 mid = lo + (hi - lo) / 2;
        if (b_rpt[mid] <= e) lo = mid;
        else hi = mid;
    }
    first[t] = lo;
}

// Workgroup (tile, ia): B nonzeros [tile * KRON_TILE, +KRON_TILE) times A
// row ia (blocks cover A rows ia0, ia0 + 1, ...).
__global__ __launch_bounds__(KRON_WG) void k_kron_fill(const u32 *__restrict__ first, u32 ntiles, u32 ia0,
                                                       const u32 *__restrict__ a_rpt, const u32 *__restrict__ a_cols,
                                                       const u32 *__restrict__ b_rpt, const u32 *__restrict__ b_cols,
                                                       u32 b_nzr, u32 b_nnz, u32 b_ncols, u32 *__restrict__ c_cols) {
    __shared__ u32 srpt[KRON_TILE + 1];
    const u32 tile = blockIdx.x % ntiles, ia = ia0 + blockIdx.x / ntiles;
    const u32 e0 = tile * KRON_TILE;
    const u32 e1 = b_nnz - e0 < KRON_TILE ? b_nnz : e0 + KRON_TILE;
    const u32 ib0 = first[tile];
    // the tile's rows start at ib0 and number at most KRON_TILE (each holds a
    // nonzero); srpt[j] = b_rpt[ib0 + j], the end row clamped to b_nzr
    const u32 nr = b_nzr - ib0 < KRON_TILE ? b_nzr - ib0 : KRON_TILE;
    for (u32 j = threadIdx.x; j <= nr; j += KRON_WG) srpt[j] = b_rpt[ib0 + j];
    __syncthreads();
    const u32 a0 = a_rpt[ia], alen = a_rpt[ia + 1] - a0;
    const u32 abase = a0 * b_nnz;
#pragma unroll
    for (u32 i = 0; i < KRON_ITEMS; ++i) {
        const u32 e = e0 + i * KRON_WG + threadIdx.x;
        if (e >= e1) break;
        u32 lo = 0, hi = nr;  // last j with srpt[j] <= e
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (srpt[mid] <= e) lo = mid;
            else hi = mid;
        }
        const u32 rs = srpt[lo], blen = srpt[lo + 1] - rs;
        const u32 bc = b_cols[e];
        u32 *out = c_cols + abase + alen * rs + (e - rs);
        for (u32 ka = 0; ka < alen; ++ka) out[ka * blen] = a_cols[a0 + ka] * b_ncols + bc;
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
                                                                         u32p(b->rows), (u32)c_nzr, b->nzr, b->nnz,
                                                                         b->nrows, (u32)c_nnz, u32p(r.rpt),
                                                                         u32p(r.rows));
            CLB_LAUNCHED("count_nnz_per_row");
            const u32 ntiles = div_up(b->nnz, KRON_TILE);
            DevArray<u32> first(ctl, ntiles);
            k_kron_tiles<<<div_up(ntiles, 256), 256, 0, ctl->stream>>>(u32p(b->rpt), b->nzr, ntiles, first.ptr);
            CLB_LAUNCHED("kron_tiles");
            // launches of at most 2^22 workgroups, whole A rows each
            const u32 rows_per = ntiles >= (1u << 22) ? 1u : (1u << 22) / ntiles;
            for (u32 ia0 = 0; ia0 < a->nzr; ia0 += rows_per) {
                const u32 nr = a->nzr - ia0 < rows_per ? a->nzr - ia0 : rows_per;
                k_kron_fill<<<ntiles * nr, KRON_WG, 0, ctl->stream>>>(first.ptr, ntiles, ia0, u32p(a->rpt),
                                                                       u32p(a->cols), u32p(b->rpt), u32p(b->cols),
                                                                       b->nzr, b->nnz, b->ncols, u32p(r.cols));
            }
            CLB_LAUNCHED("calculate_kronecker_product");
            sync(ctl);
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
