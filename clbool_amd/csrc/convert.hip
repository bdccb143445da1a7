// convert.hip -- format conversions between COO, CSR and DCSR.
//
// Restates src/common/matrices_conversions.cpp: coo_to_dcsr_shallow (:78-91,
// rows compressed by flag/scan/scatter, cols shared), dcsr_to_coo_shallow /
// _deep (:43-76), dcsr_to_csr (:247-271, cols shared, nnz checked) and
// csr_to_dcsr (:211-244, empty rows dropped, cols shared).
#include "sort.hpp"

#include "device_utils.hpp"

namespace clb {

// prepare_array_for_rows_positions (prepare_positions.cl:29-44), with a
// trailing 0 so that an exclusive scan of n+1 entries yields nzr at [n].
__global__ void k_row_start_flags(const u32 *__restrict__ rows, u32 n, u32 *__restrict__ flags) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || rows[i] != rows[i - 1]) ? 1u : 0u;
    if (i == n) flags[n] = 0;
}

// set_positions_rows (set_positions.cl:89-113)
__global__ void k_set_rows(const u32 *__restrict__ rows, const u32 *__restrict__ pos, u32 n, u32 nzr,
                           u32 *__restrict__ rpt, u32 *__restrict__ crows) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) rpt[nzr] = n;
    if (i < n && pos[i] != pos[i + 1]) {
        rpt[pos[i]] = i;
        crows[pos[i]] = rows[i];
    }
}

void rows_to_dcsr(clb_controls_s *ctl, const u32 *rows, u32 nnz, clb_buffer_s **rpt_out, clb_buffer_s **rows_out,
                  u32 *nzr_out) {
    DevArray<u32> pos(ctl, (u64)nnz + 1);
    k_row_start_flags<<<div_up((u64)nnz + 1, 256), 256, 0, ctl->stream>>>(rows, nnz, pos.ptr);
    CLB_LAUNCHED("prepare_array_for_rows_positions");
    DevArray<u32> tot(ctl, 1);
    exclusive_scan_u32(ctl, pos.ptr, pos.ptr, (u64)nnz + 1, tot.ptr);
    u32 nzr = 0;
    readback(ctl, &nzr, tot.ptr, 4);
    clb_buffer_s *rpt = buffer_new(ctl, 4ull * (nzr + 1ull));
    clb_buffer_s *cr = nullptr;
    try {
        cr = buffer_new(ctl, 4ull * (nzr ? nzr : 1));
    } catch (...) {
        buffer_release(rpt);
        throw;
    }
    k_set_rows<<<div_up(nnz, 256), 256, 0, ctl->stream>>>(rows, pos.ptr, nnz, nzr, u32p(rpt), u32p(cr));
    CLB_LAUNCHED("set_positions_rows");
    sync(ctl);
    *rpt_out = rpt;
    *rows_out = cr;
    *nzr_out = nzr;
}

// dscr_to_coo (dscr_to_coo.cl:8-23): one wave per stored row.
__global__ __launch_bounds__(256) void k_expand_rows(const u32 *__restrict__ rpt, const u32 *__restrict__ rows,
                                                     u32 nzr, u32 *__restrict__ out) {
    const u32 r = blockIdx.x * 4 + dev::wave_id();
    if (r >= nzr) return;
    const u32 v = rows[r], e = rpt[r + 1];
    for (u32 j = rpt[r] + dev::lane(); j < e; j += 64) out[j] = v;
}

void dcsr_expand_rows(clb_controls_s *ctl, const u32 *rpt, const u32 *rows, u32 nzr, u32 nnz, u32 *out_rows) {
    (void)nnz;
    if (!nzr) return;
    k_expand_rows<<<div_up(nzr, 4), 256, 0, ctl->stream>>>(rpt, rows, nzr, out_rows);
    CLB_LAUNCHED("dscr_to_coo");
}

// dcsr_to_csr_set_size (conversions.cl:8-18)
__global__ void k_dcsr_row_sizes(const u32 *__restrict__ rpt, const u32 *__restrict__ rows, u32 nzr, u32 nrows,
                                 u32 *__restrict__ csr_rpt) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nzr && rows[i] <= nrows) csr_rpt[rows[i]] = rpt[i + 1] - rpt[i];
}

// prepare_for_shift_empty_rows (prepare_positions.cl:69-84)
__global__ void k_nonempty_flags(const u32 *__restrict__ rpt, u32 n, u32 *__restrict__ flags) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = rpt[i] != rpt[i + 1];
    if (i == n) flags[n] = 0;
}

// set_positions_pointers_and_rows_csr (set_positions.cl:69-85)
__global__ void k_csr_set_positions(const u32 *__restrict__ rpt, const u32 *__restrict__ pos, u32 nrows,
                                    u32 *__restrict__ c_rpt, u32 *__restrict__ c_rows) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) c_rpt[pos[nrows]] = rpt[nrows];
    if (i < nrows && pos[i] != pos[i + 1]) {
        c_rpt[pos[i]] = rpt[i];
        c_rows[pos[i]] = i;
    }
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_coo_to_dcsr_shallow(clb_controls ctl, const clb_coo *a, clb_dcsr *out) {
    return run_guarded([&] {
        clb_dcsr r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz != 0) {
            rows_to_dcsr(ctl, u32p(a->rows), a->nnz, &r.rpt, &r.rows, &r.nzr);
            buffer_retain(a->cols);  // shallow: the columns are shared
            r.cols = a->cols;
            r.nnz = a->nnz;
        }
        *out = r;
    });
}

static void dcsr_to_coo(clb_controls ctl, const clb_dcsr *a, clb_coo *out, bool deep) {
    clb_coo r{};
    r.nrows = a->nrows;
    r.ncols = a->ncols;
    if (a->nnz != 0) {
        r.rows = buffer_new(ctl, 4ull * a->nnz);
        try {
            dcsr_expand_rows(ctl, u32p(a->rpt), u32p(a->rows), a->nzr, a->nnz, u32p(r.rows));
            if (deep) {
                r.cols = buffer_new(ctl, 4ull * a->nnz);
                CLB_HIP(hipMemcpyAsync(r.cols->ptr, a->cols->ptr, 4ull * a->nnz, hipMemcpyDeviceToDevice,
                                       ctl->stream),
                        COPY_BUFFER_ERROR);
            } else {
                buffer_retain(a->cols);
                r.cols = a->cols;
            }
            sync(ctl);
        } catch (...) {
            clb_coo_release(&r);
            throw;
        }
        r.nnz = a->nnz;
    }
    *out = r;
}

clb_status clb_dcsr_to_coo_shallow(clb_controls ctl, const clb_dcsr *a, clb_coo *out) {
    return run_guarded([&] { dcsr_to_coo(ctl, a, out, false); });
}

clb_status clb_dcsr_to_coo_deep(clb_controls ctl, const clb_dcsr *a, clb_coo *out) {
    return run_guarded([&] { dcsr_to_coo(ctl, a, out, true); });
}

clb_status clb_dcsr_to_csr(clb_controls ctl, const clb_dcsr *a, clb_csr *out) {
    return run_guarded([&] {
        clb_csr r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz == 0) {  // :248-250
            *out = r;
            return;
        }
        const u64 n1 = (u64)a->nrows + 1;
        r.rpt = buffer_new(ctl, 4ull * n1);
        try {
            fill_u32(ctl, u32p(r.rpt), n1, 0);
            k_dcsr_row_sizes<<<div_up(a->nzr, 256), 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(a->rows), a->nzr,
                                                                             a->nrows, u32p(r.rpt));
            CLB_LAUNCHED("dcsr_to_csr_set_size");
            DevArray<u32> tot(ctl, 1);
            exclusive_scan_u32(ctl, u32p(r.rpt), u32p(r.rpt), n1, tot.ptr);
            u32 c_nnz = 0;
            readback(ctl, &c_nnz, tot.ptr, 4);
            // :268
            if (c_nnz != a->nnz) raise(INVALID_VALUE, "nnz in dcsr and csr matrices should be equal!");
        } catch (...) {
            buffer_release(r.rpt);
            throw;
        }
        buffer_retain(a->cols);
        r.cols = a->cols;
        r.nnz = a->nnz;
        *out = r;
    });
}

clb_status clb_csr_to_dcsr(clb_controls ctl, const clb_csr *a, clb_dcsr *out) {
    return run_guarded([&] {
        clb_dcsr r{};
        r.nrows = a->nrows;
        r.ncols = a->ncols;
        if (a->nnz == 0) {  // :212-214
            *out = r;
            return;
        }
        const u64 n1 = (u64)a->nrows + 1;
        DevArray<u32> pos(ctl, n1);
        k_nonempty_flags<<<div_up(n1, 256), 256, 0, ctl->stream>>>(u32p(a->rpt), a->nrows, pos.ptr);
        CLB_LAUNCHED("prepare_for_shift_empty_rows");
        DevArray<u32> tot(ctl, 1);
        exclusive_scan_u32(ctl, pos.ptr, pos.ptr, n1, tot.ptr);
        u32 nzr = 0;
        readback(ctl, &nzr, tot.ptr, 4);
        r.rpt = buffer_new(ctl, 4ull * (nzr + 1ull));
        try {
            r.rows = buffer_new(ctl, 4ull * (nzr ? nzr : 1));
            k_csr_set_positions<<<div_up(n1, 256), 256, 0, ctl->stream>>>(u32p(a->rpt), pos.ptr, a->nrows,
                                                                           u32p(r.rpt), u32p(r.rows));
            CLB_LAUNCHED("set_positions_pointers_and_rows_csr");
            sync(ctl);
        } catch (...) {
            clb_dcsr_release(&r);
            throw;
        }
        buffer_retain(a->cols);
        r.cols = a->cols;
        r.nnz = a->nnz;
        r.nzr = nzr;
        *out = r;
    });
}

}  // extern "C"
