// esc.hip -- the expand-sort-compress SpGEMM, dcsr::matrix_multiplication
// (src/dcsr/dcsr_matrix_multiplication.cpp:26-168).
//
// A second GPU SpGEMM that shares no accumulator with the hash path, so the
// two cross-check each other.  The reference sizes every row by its product
// count (count_workload, :388-404; "pre" matrix at the products prefix,
// :279-334), then writes each row's sorted unique columns into its slot with
// a kernel per bin (copy_one_value for one product, heap_merge up to 32,
// bitonic_esc up to 512, merge_large_rows above, :192-275) and compacts
// (create_final_matrix, :97-158).  Here, with the same one-pass structure:
//  * rows of at most 64 products: one wave, register bitonic sort (k_tiny);
//  * up to 2^13 products: one workgroup expands the products into LDS,
//    merge-sorts them there and keeps the first of each run (k_esc);
//  * longer rows: all their products are expanded at once into 64-bit keys
//    (row << 32 | col) in HBM, radix-sorted over the bits the shape needs and
//    deduplicated straight into a DCSR (the COO-construction machinery),
//    instead of the reference's pairwise merges of B rows per row;
//  * exact sizes are scanned and every row is copied to its final place.
#include "sort.hpp"
#include "spgemm.hpp"

namespace clb {
namespace spg {

// Slots of the rows sorted in LDS or in registers: their product count;
// rows sorted in HBM get 0 here and their product count in klen.
__global__ void k_esc_sizes(const u32 *__restrict__ prod, u32 n, u32 lds_max, u32 *__restrict__ tlen,
                            u32 *__restrict__ klen) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32 P = prod[i];
    tlen[i] = P <= lds_max ? P : 0u;
    klen[i] = P > lds_max ? P : 0u;
}

// Expands the products of each long row into keys (row << 32 | col) at
// koff[row]; one workgroup per row, slots taken by wave-aggregated appends.
constexpr int EXP_WG = 256;
__global__ __launch_bounds__(EXP_WG) void k_esc_expand(Mats g, const u32 *__restrict__ perm,
                                                       const u64 *__restrict__ koff, u64 *__restrict__ keys) {
    __shared__ __align__(8) u32 big[BIG_WORDS(EXP_WG)];
    __shared__ u32 cnt;
    const u32 r = perm[blockIdx.x];
    const u32 i = g.row0 + r;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    u64 *dst = keys + koff[r];
    const u64 hi = (u64)r << 32;
    auto append = [&](u32 col) {
        const unsigned long long m = __ballot(1);
        const u32 leader = (u32)__ffsll(m) - 1u;
        u32 base = 0;
        if ((u32)dev::lane() == leader) base = atomicAdd(&cnt, (u32)__popcll(m));
        base = __shfl(base, (int)leader, dev::WAVE);
        dst[base + dev::lanes_below(m)] = hi | col;
    };
    for_each_product<EXP_WG, 8, BIG_ROW>(g, g.a_rpt[i], g.a_rpt[i + 1], big, append);
}

// Sizes and sources of the long rows from their deduplicated DCSR L (row ids
// are block rows): nnz_row, the flag that the row comes from L, its start.
__global__ void k_esc_long_rows(const u32 *__restrict__ l_rpt, const u32 *__restrict__ l_rows, u32 l_nzr,
                                u32 *__restrict__ nnz_row, u32 *__restrict__ from_side, u64 *__restrict__ soff) {
    const u32 q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= l_nzr) return;
    const u32 r = l_rows[q];
    nnz_row[r] = l_rpt[q + 1] - l_rpt[q];
    from_side[r] = 1;
    soff[r] = l_rpt[q];
}

static int bits_of(u64 maxv) {
    int b = 0;
    while (b < 32 && (1ull << b) <= maxv) ++b;
    return b;
}

void spgemm_esc(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, clb_dcsr &out, clb_spgemm_stats &st) {
    std::memset(&st, 0, sizeof(st));
    check_dims(a, b);
    out = clb_dcsr{};
    out.nrows = a.nrows;
    out.ncols = b.ncols;
    // dcsr_matrix_multiplication.cpp:37-40: an empty operand gives an empty result
    if (a.nnz == 0 || b.nnz == 0 || a.nzr == 0) return;
    const u32 nrows = a.nzr;
    Setup s;
    prepare(ctl, a, b, 0, nrows, s, true);
    s.cfg.esc_only = 1;
    s.cfg.tiny_max = 64;
    const u32 lds_max = 1u << ESC_MAX_BITS;
    DevArray<u32> nnz_row(ctl, nrows), tlen(ctl, nrows), klen(ctl, nrows), from_side(ctl, nrows);
    DevArray<u64> toff(ctl, (u64)nrows + 1), koff(ctl, (u64)nrows + 1), soff(ctl, nrows);
    k_esc_sizes<<<div_up(nrows, 256), 256, 0, ctl->stream>>>(s.P, nrows, lds_max, tlen.ptr, klen.ptr);
    CLB_LAUNCHED("esc_sizes");
    exclusive_scan_u32_to_u64(ctl, tlen.ptr, toff.ptr, nrows, toff.ptr + nrows);
    exclusive_scan_u32_to_u64(ctl, klen.ptr, koff.ptr, nrows, koff.ptr + nrows);
    fill_u32(ctl, from_side.ptr, nrows, 0u);
    DevArray<u32> perm;
    BinBuf bb;
    BinHost h;
    bin_rows(ctl, s, nullptr, nnz_row.ptr, perm, bb);
    read_bins(ctl, bb, h);
    u64 tot[3];
    readback(ctl, &tot[0], s.scal.ptr, 8);
    readback(ctl, &tot[1], toff.ptr + nrows, 8);
    readback(ctl, &tot[2], koff.ptr + nrows, 8);
    st.products = tot[0];
    st.rows = nrows;
    st.n_bins = NBINS;
    if (tot[0] == 0) return;
    // rows sorted in registers or in LDS: one pass into slots of their
    // product count (never too small)
    DevArray<u32> temp(ctl, tot[1] ? tot[1] : 1);
    Out o{nnz_row.ptr, toff.ptr, temp.ptr, s.P, reinterpret_cast<u32 *>(s.scal.ptr + 5), Dump{}, 0};
    launch_bins(ctl, s, perm.ptr, h.off, ONE, o, nullptr, CLB_KP_ONEPASS);
    // long rows: expand, sort and deduplicate in HBM
    DevArray<u64> keys;
    clb_dcsr L{};
    const u32 nlong = h.off[NBINS] - h.off[BIN_BITMAP0];
    try {
        if (nlong) {
            keys.reset(ctl, tot[2]);
            k_esc_expand<<<nlong, EXP_WG, 0, ctl->stream>>>(s.g, perm.ptr + h.off[BIN_BITMAP0], koff.ptr, keys.ptr);
            CLB_LAUNCHED("esc_expand");
            radix_sort_keys(ctl, keys.ptr, tot[2], bits_of(b.ncols ? b.ncols - 1 : 0), bits_of(nrows - 1));
            L.nrows = nrows;
            L.ncols = b.ncols;
            keys_to_dcsr(ctl, keys.ptr, tot[2], L);
            keys.free_();
            if (L.nzr) {
                k_esc_long_rows<<<div_up(L.nzr, 256), 256, 0, ctl->stream>>>(
                    u32p(L.rpt), u32p(L.rows), L.nzr, nnz_row.ptr, from_side.ptr, soff.ptr);
                CLB_LAUNCHED("esc_long_rows");
            }
        }
        u64 fails = 0;
        readback(ctl, &fails, s.scal.ptr + 5, 4);
        if ((u32)fails) raise(RUN_KERNEL_FAILURE, "ESC row overflowed its product-count slot");
        DevArray<u64> offs;
        DevArray<u32> pos;
        scan_sizes(ctl, s, nnz_row.ptr, offs, pos);
        u64 sc[3];
        readback(ctl, sc, s.scal.ptr, sizeof(sc));
        const u64 nnz = sc[1];
        const u32 nzr = static_cast<u32>(sc[2]);
        if (nnz > 0xFFFFFFFFull) raise(INVALID_VALUE, "result has more nonzeros than uint32 indices allow");
        if (nnz) {
            clb_buffer_s *cols = buffer_new(ctl, 4ull * nnz);
            try {
                compact_rows(ctl->stream, nrows, nnz, nnz_row.ptr, offs.ptr, temp.ptr, toff.ptr, u32p(L.cols),
                             soff.ptr, from_side.ptr, nullptr, nullptr, u32p(cols), nullptr, true);
                assemble(ctl, s, a, nnz_row.ptr, offs.ptr, pos.ptr, nnz, nzr, cols, out);
                sync(ctl);
            } catch (...) {
                (void)hipStreamSynchronize(ctl->stream);
                clb_dcsr_release(&out);
                if (!out.cols) buffer_release(cols);
                throw;
            }
        }
        st.nnz = nnz;
        st.nzr = nzr;
    } catch (...) {
        (void)hipStreamSynchronize(ctl->stream);
        clb_dcsr_release(&L);
        throw;
    }
    clb_dcsr_release(&L);
}

}  // namespace spg
}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_dcsr_spgemm_esc(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out) {
    return run_guarded([&] {
        if (!a || !b || !out) raise(INVALID_ARGUMENT, "null argument");
        clb_spgemm_stats st;
        clb_dcsr r{};
        spg::spgemm_esc(ctl, *a, *b, r, st);
        ctl->last = st;
        *out = r;
    });
}

}  // extern "C"
