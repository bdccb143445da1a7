// dcsr_ops.hip -- dcsr::transpose, dcsr::submatrix and dcsr::reduce.
//
//  * transpose (src/dcsr/dcsr_transpose.cpp:5-17): the reference expands the
//    matrix to COO, swaps the row and column arrays and re-sorts them through
//    the matrix_coo constructor.  Here the (col, row) pairs are packed into
//    64-bit keys and radix-sorted over only the bits the shape needs, then
//    split into the transposed DCSR.
//  * submatrix (src/dcsr/dcsr_submatrix.cpp:98-156, submatrix.cl): rows
//    [i, i+nrows) and columns [j, j+ncols), shifted to the origin.  Row range
//    and per-row column ranges come from lower bounds; empty sub-rows are
//    dropped.  Bounds are computed in 64 bits (the reference's i + nrows and
//    j + ncols wrap in uint32).
//  * reduce (src/dcsr/dcsr_reduce.cpp:5-31, reduce.cl): OR over columns: one
//    entry (column 0) per nonempty row.
//  * digest (no counterpart; it replaces reading a matrix back to compare it
//    on the host, as compare_matrices does, src/common/utils.cpp:42-71): a
//    64-bit sum over the nonzeros of a mix of (position, row id, column).
#include "clb_internal.hpp"

#include "sort.hpp"

namespace clb {

namespace {

__device__ __forceinline__ u32 lower_bound_u32(const u32 *__restrict__ a, u32 n, u64 v) {
    u32 lo = 0, hi = n;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo) / 2;
        if ((u64)a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// rows_range (submatrix.cl:38-49): out[0] = lower_bound(rows, i),
// out[1] = lower_bound(rows, i + nrows).
__global__ void k_rows_range(const u32 *__restrict__ rows, u32 nzr, u64 v0, u64 v1, u32 *__restrict__ out) {
    const u32 t = threadIdx.x;
    if (t < 2) out[t] = lower_bound_u32(rows, nzr, t == 0 ? v0 : v1);
}

// submatrix_count_nnz (submatrix.cl:56-83): per row of the range, the start
// and length of its column sub-range [j, jend).
__global__ void k_sub_count(const u32 *__restrict__ rpt, const u32 *__restrict__ cols, u32 rb, u32 n, u64 j,
                            u64 jend, u32 *__restrict__ cnt, u32 *__restrict__ start) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const u32 s = rpt[rb + t], e = rpt[rb + t + 1];
    const u32 lo = lower_bound_u32(cols + s, e - s, j);
    const u32 hi = lower_bound_u32(cols + s, e - s, jend);
    cnt[t] = hi - lo;
    start[t] = s + lo;
}

// submatrix_fill_nnz (submatrix.cl:86-109): one wave per row copies its
// sub-row shifted by j.
__global__ __launch_bounds__(256) void k_sub_fill(const u32 *__restrict__ cnt, const u32 *__restrict__ start,
                                                  const u32 *__restrict__ off, u32 n, const u32 *__restrict__ cols,
                                                  u32 j, u32 *__restrict__ out) {
    const u32 t = blockIdx.x * 4 + threadIdx.x / 64;
    if (t >= n) return;
    const u32 len = cnt[t], s = start[t], o = off[t];
    for (u32 k = threadIdx.x % 64; k < len; k += 64) out[o + k] = cols[s + k] - j;
}

// set_positions_with_offset (set_positions.cl:115-139): keep the nonempty
// sub-rows, rows shifted by i.
__global__ void k_sub_rows(const u32 *__restrict__ cnt, const u32 *__restrict__ off, const u32 *__restrict__ pos,
                           u32 n, const u32 *__restrict__ rows, u32 rb, u32 i, u32 *__restrict__ c_rpt,
                           u32 *__restrict__ c_rows, u32 nnz, u32 nzr) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) c_rpt[nzr] = nnz;
    if (t >= n || cnt[t] == 0) return;
    c_rpt[pos[t]] = off[t];
    c_rows[pos[t]] = rows[rb + t] - i;
}

__global__ void k_nonzero(const u32 *__restrict__ cnt, u32 n, u32 *__restrict__ flags) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) flags[t] = cnt[t] != 0;
}

// set_rpt_and_cols (reduce.cl:11-23).
__global__ void k_reduce_rpt(u32 *__restrict__ rpt, u32 *__restrict__ cols, u32 n) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        cols[t] = 0;
        rpt[t] = t;
    }
    if (t == 0) rpt[n] = n;
}

void release_dcsr(clb_dcsr &m) {
    buffer_release(m.rpt);
    buffer_release(m.rows);
    buffer_release(m.cols);
    m = clb_dcsr{};
}

// splitmix64's finaliser (Steele, Lea and Flood 2014).
__device__ __forceinline__ u64 mix64(u64 x) {
    u64 z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Sum over stored rows r and their positions q of
// mix64(mix64(q) ^ (row << 32 | col)); one wave per row (grid-stride), one
// atomic per workgroup.
__global__ __launch_bounds__(256) void k_dcsr_digest(const u32 *__restrict__ rpt, const u32 *__restrict__ rows,
                                                     const u32 *__restrict__ cols, u32 nzr,
                                                     unsigned long long *__restrict__ out) {
    __shared__ u64 part[4];
    u64 acc = 0;
    const u32 w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (u32 r = blockIdx.x * 4 + w; r < nzr; r += gridDim.x * 4) {
        const u64 hi = (u64)rows[r] << 32;
        const u32 e = rpt[r + 1];
        for (u32 q = rpt[r] + l; q < e; q += 64) acc += mix64(mix64(q) ^ (hi | cols[q]));
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (l == 0) part[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

}  // namespace

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_dcsr_digest(clb_controls ctl, const clb_dcsr *a, uint64_t *digest) {
    return run_guarded([&] {
        if (!a || !digest) raise(INVALID_ARGUMENT, "null argument");
        u64 h[2] = {0, 0};
        if (a->nzr && a->nnz) {
            DevArray<u64> acc(ctl, 1);
            fill_u64(ctl, acc.ptr, 1, 0ull);
            const u32 grid = div_up(a->nzr, 4) < 8192u ? div_up(a->nzr, 4) : 8192u;
            k_dcsr_digest<<<grid, 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(a->rows), u32p(a->cols), a->nzr,
                                                         reinterpret_cast<unsigned long long *>(acc.ptr));
            CLB_LAUNCHED("dcsr_digest");
            readback(ctl, h, acc.ptr, 8);
        }
        // the shape and sizes take part too
        *digest = h[0] ^ (((u64)a->nnz << 32 | a->nzr) * 0x9E3779B97F4A7C15ull) ^
                  ((u64)a->nrows << 32 | a->ncols);
    });
}

clb_status clb_dcsr_transpose(clb_controls ctl, const clb_dcsr *a, clb_dcsr *out) {
    return run_guarded([&] {
        clb_dcsr r{};
        r.nrows = a->ncols;
        r.ncols = a->nrows;
        if (a->nnz == 0) {  // :6-9
            *out = r;
            return;
        }
        const u64 n = a->nnz;
        DevArray<u32> arow(ctl, n);
        dcsr_expand_rows(ctl, u32p(a->rpt), u32p(a->rows), a->nzr, a->nnz, arow.ptr);
        DevArray<u64> keys(ctl, n);
        // key = col << 32 | row: sorted, it is the transposed row-major order
        pack_pairs(ctl, u32p(a->cols), arow.ptr, n, keys.ptr);
        sort_pair_keys(ctl, keys.ptr, n, a->ncols ? a->ncols - 1 : 0, a->nrows ? a->nrows - 1 : 0);
        clb_buffer_s *rows = buffer_new(ctl, 4 * n);
        try {
            r.cols = buffer_new(ctl, 4 * n);
            unpack_pairs(ctl, keys.ptr, n, u32p(rows), u32p(r.cols));
            rows_to_dcsr(ctl, u32p(rows), a->nnz, &r.rpt, &r.rows, &r.nzr);
            r.nnz = a->nnz;
            buffer_release(rows);
        } catch (...) {
            buffer_release(rows);
            release_dcsr(r);
            throw;
        }
        *out = r;
    });
}

clb_status clb_dcsr_submatrix(clb_controls ctl, const clb_dcsr *a, uint32_t i, uint32_t j, uint32_t nrows,
                              uint32_t ncols, clb_dcsr *out) {
    return run_guarded([&] {
        // every empty outcome is a 0 x 0 matrix, as in the reference
        // (:101-108, :112-115, :125-128)
        clb_dcsr r{};
        if (a->nnz == 0 || a->nzr == 0) {
            *out = r;
            return;
        }
        DevArray<u32> range(ctl, 2);
        k_rows_range<<<1, 64, 0, ctl->stream>>>(u32p(a->rows), a->nzr, i, (u64)i + nrows, range.ptr);
        CLB_LAUNCHED("rows_range");
        u32 rr[2];
        readback(ctl, rr, range.ptr, sizeof(rr));
        const u32 rb = rr[0], n = rr[1] - rr[0];
        if (n == 0) {
            *out = r;
            return;
        }
        DevArray<u32> cnt(ctl, n), start(ctl, n), off(ctl, n), pos(ctl, n), tot(ctl, 2);
        k_sub_count<<<div_up(n, 256), 256, 0, ctl->stream>>>(u32p(a->rpt), u32p(a->cols), rb, n, j, (u64)j + ncols,
                                                             cnt.ptr, start.ptr);
        CLB_LAUNCHED("submatrix_count_nnz");
        exclusive_scan_u32(ctl, cnt.ptr, off.ptr, n, tot.ptr);
        k_nonzero<<<div_up(n, 256), 256, 0, ctl->stream>>>(cnt.ptr, n, pos.ptr);
        CLB_LAUNCHED("prepare_for_shift_empty_rows");
        exclusive_scan_u32(ctl, pos.ptr, pos.ptr, n, tot.ptr + 1);
        u32 t[2];
        readback(ctl, t, tot.ptr, sizeof(t));
        const u32 nnz = t[0], nzr = t[1];
        if (nnz == 0) {
            *out = r;
            return;
        }
        r.nrows = nrows;
        r.ncols = ncols;
        try {
            r.cols = buffer_new(ctl, 4ull * nnz);
            r.rpt = buffer_new(ctl, 4ull * (nzr + 1));
            r.rows = buffer_new(ctl, 4ull * nzr);
            k_sub_fill<<<div_up(n, 4), 256, 0, ctl->stream>>>(cnt.ptr, start.ptr, off.ptr, n, u32p(a->cols), j,
                                                              u32p(r.cols));
            CLB_LAUNCHED("submatrix_fill_nnz");
            k_sub_rows<<<div_up(n, 256), 256, 0, ctl->stream>>>(cnt.ptr, off.ptr, pos.ptr, n, u32p(a->rows), rb, i,
                                                                u32p(r.rpt), u32p(r.rows), nnz, nzr);
            CLB_LAUNCHED("set_positions_with_offset");
            r.nnz = nnz;
            r.nzr = nzr;
            sync(ctl);
        } catch (...) {
            release_dcsr(r);
            throw;
        }
        *out = r;
    });
}

clb_status clb_dcsr_reduce(clb_controls ctl, const clb_dcsr *a, clb_dcsr *out) {
    return run_guarded([&] {
        clb_dcsr r{};
        if (a->nnz == 0) {  // :7-10
            *out = r;
            return;
        }
        const u32 n = a->nzr;
        r.nrows = a->nrows;
        r.ncols = 1;
        try {
            r.rpt = buffer_new(ctl, 4ull * (n + 1));
            r.rows = buffer_new(ctl, 4ull * n);
            r.cols = buffer_new(ctl, 4ull * n);
            CLB_HIP(hipMemcpyAsync(r.rows->ptr, a->rows->ptr, 4ull * n, hipMemcpyDeviceToDevice, ctl->stream),
                    COPY_BUFFER_ERROR);
            k_reduce_rpt<<<div_up(n + 1, 256), 256, 0, ctl->stream>>>(u32p(r.rpt), u32p(r.cols), n);
            CLB_LAUNCHED("set_rpt_and_cols");
            r.nnz = n;
            r.nzr = n;
            sync(ctl);
        } catch (...) {
            release_dcsr(r);
            throw;
        }
        *out = r;
    });
}

}  // extern "C"
