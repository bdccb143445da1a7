// closure.hip -- batched closure step of CFPQ/RPQ evaluation: X_t |= Y_t * Z_t
// for a set of triples t over label matrices that share one vertex set.
//
// clBool itself has no closure driver: its user (spbla) calls
// dcsr::matrix_multiplication_hash and an addition once per triple.  The
// per-triple loop is kept here as the baseline (mode CLB_CLOSURE_PER_TRIPLE).
// The batched mode runs all products of a step as ONE SpGEMM on stacked
// operands: Y' = [Y_0 | Y_1 | ...] (row block t holds Y_t, its columns shifted
// by t * n) times Z' = [Z_0; Z_1; ...] (row block t holds Z_t) gives row block
// t of the result = Y_t * Z_t, so every kernel launch of the SpGEMM covers the
// rows of all products.  The unions are one sort and deduplication of the
// keys of all X blocks and all products, followed by a split back into the
// label matrices.  Every step of both modes is bit-exact: the products and
// unions are exact set operations, only the launch structure differs.
#include "sort.hpp"
#include "spgemm.hpp"

#include <algorithm>
#include <chrono>

namespace clb {

// One matrix of a stacked operand.
struct Seg {
    const u32 *rpt, *rows, *cols;
    u32 nzr, nnz;
    u32 row_shift, col_shift;  // added to row ids and to column ids
    u32 nzr_off, nnz_off;      // position of the block in the stack
};

constexpr int MAXSEG = 256;

// Copies the blocks' rows (ids shifted) and row pointers (offset by the
// block's nonzero position) into the stack; one thread per stored row.
__global__ void k_stack_rows(const Seg *__restrict__ segs, u32 nseg, u32 total_nzr, u32 *__restrict__ rpt,
                             u32 *__restrict__ rows, u32 total_nnz) {
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) rpt[total_nzr] = total_nnz;
    if (g >= total_nzr) return;
    u32 lo = 0, hi = nseg;  // last segment with nzr_off <= g
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) / 2;
        if (segs[mid].nzr_off <= g) lo = mid;
        else hi = mid;
    }
    while (lo + 1 < nseg && segs[lo + 1].nzr_off <= g) ++lo;  // skip empty segments
    const Seg &s = segs[lo];
    const u32 i = g - s.nzr_off;
    rows[g] = s.rows[i] + s.row_shift;
    rpt[g] = s.rpt[i] + s.nnz_off;
}

// Copies the blocks' columns (shifted); one thread per nonzero.
__global__ void k_stack_cols(const Seg *__restrict__ segs, u32 nseg, u32 total_nnz, u32 *__restrict__ cols) {
    const u32 e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total_nnz) return;
    u32 lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) / 2;
        if (segs[mid].nnz_off <= e) lo = mid;
        else hi = mid;
    }
    while (lo + 1 < nseg && segs[lo + 1].nnz_off <= e) ++lo;
    const Seg &s = segs[lo];
    cols[e] = s.cols[e - s.nnz_off] + s.col_shift;
}

// Keys (row << 32 | col) of a DCSR matrix with rows shifted, one wave per row.
__global__ __launch_bounds__(256) void k_keys_shifted(const u32 *__restrict__ rpt, const u32 *__restrict__ rows,
                                                      const u32 *__restrict__ cols, u32 nzr, u32 row_shift,
                                                      u64 *__restrict__ keys) {
    const u32 r = blockIdx.x * 4 + dev::wave_id();
    if (r >= nzr) return;
    const u64 hi = (u64)(rows[r] + row_shift) << 32;
    const u32 e = rpt[r + 1];
    for (u32 j = rpt[r] + dev::lane(); j < e; j += 64) keys[j] = hi | cols[j];
}

// First key index of every row block [b*n, (b+1)*n) of sorted unique keys.
__global__ void k_block_bounds(const u64 *__restrict__ keys, u64 m, u32 n, u32 nblocks, u64 *__restrict__ bound) {
    const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblocks) return;
    const u64 target = (u64)b * n << 32;
    u64 lo = 0, hi = m;
    while (lo < hi) {
        const u64 mid = (lo + hi) / 2;
        if (keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    bound[b] = lo;
}

// Block b of sorted unique keys -> (rows - b*n, cols) arrays of that block.
__global__ void k_unstack_pairs(const u64 *__restrict__ keys, u64 k0, u32 cnt, u32 row_shift, u32 *__restrict__ rows,
                                u32 *__restrict__ cols) {
    const u32 e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cnt) return;
    const u64 k = keys[k0 + e];
    rows[e] = (u32)(k >> 32) - row_shift;
    cols[e] = (u32)k;
}

// Builds the stack of the given matrices on the device.
static clb_dcsr stack(clb_controls_s *ctl, const std::vector<Seg> &segs_in, u32 nrows, u32 ncols) {
    std::vector<Seg> segs = segs_in;
    u64 nzr = 0, nnz = 0;
    for (auto &s : segs) {
        s.nzr_off = (u32)nzr;
        s.nnz_off = (u32)nnz;
        nzr += s.nzr;
        nnz += s.nnz;
    }
    if (nnz > 0xFFFFFFFFull) raise(INVALID_VALUE, "stacked operand exceeds uint32 nonzeros");
    clb_dcsr m{};
    m.nrows = nrows;
    m.ncols = ncols;
    if (!nnz) return m;
    DevArray<Seg> dseg(ctl, segs.size());
    CLB_HIP(hipMemcpyAsync(dseg.ptr, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice, ctl->stream),
            WRITE_BUFFER_ERROR);
    m.rpt = buffer_new(ctl, 4 * (nzr + 1));
    m.rows = buffer_new(ctl, 4 * nzr);
    m.cols = buffer_new(ctl, 4 * nnz);
    k_stack_rows<<<div_up(nzr + 1, 256), 256, 0, ctl->stream>>>(dseg.ptr, (u32)segs.size(), (u32)nzr, u32p(m.rpt),
                                                                 u32p(m.rows), (u32)nnz);
    CLB_LAUNCHED("stack_rows");
    k_stack_cols<<<div_up(nnz, 256), 256, 0, ctl->stream>>>(dseg.ptr, (u32)segs.size(), (u32)nnz, u32p(m.cols));
    CLB_LAUNCHED("stack_cols");
    sync(ctl);  // dseg is released on return
    m.nnz = (u32)nnz;
    m.nzr = (u32)nzr;
    return m;
}

static Seg seg_of(const clb_dcsr &m, u32 row_shift, u32 col_shift) {
    return Seg{u32p(m.rpt), u32p(m.rows), u32p(m.cols), m.nzr, m.nnz, row_shift, col_shift, 0, 0};
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_closure_step(clb_controls ctl, clb_dcsr *mats, uint32_t nmats, const uint32_t *triples,
                            uint32_t ntriples, const clb_closure_opts *opts, clb_closure_stats *stats) {
    return run_guarded([&] {
        clb_closure_opts o{};
        if (opts) o = *opts;
        clb_closure_stats st{};
        auto t0 = std::chrono::steady_clock::now();
        if (!nmats) raise(INVALID_ARGUMENT, "no matrices");
        const u32 n = mats[0].nrows;
        for (u32 k = 0; k < nmats; ++k)
            if (mats[k].nrows != n || mats[k].ncols != n)
                raise(INVALID_ARGUMENT, "closure matrices must share one square vertex set");
        for (u32 t = 0; t < 3 * ntriples; ++t)
            if (triples[t] >= nmats) raise(INVALID_ARGUMENT, "triple names an unknown matrix");
        std::vector<u64> before(nmats);
        for (u32 k = 0; k < nmats; ++k) before[k] = mats[k].nnz;
        // distinct X matrices of the step
        std::vector<u32> xs;
        for (u32 t = 0; t < ntriples; ++t) xs.push_back(triples[3 * t]);
        std::sort(xs.begin(), xs.end());
        xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
        auto xid = [&](u32 m) { return (u32)(std::lower_bound(xs.begin(), xs.end(), m) - xs.begin()); };
        clb_spgemm_opts so{};
        so.mode = o.spgemm_mode;

        if (o.mode == CLB_CLOSURE_PER_TRIPLE) {
            // the reference-style loop: products from the step's start values,
            // then one union per triple
            std::vector<clb_dcsr> prods(ntriples);
            try {
                for (u32 t = 0; t < ntriples; ++t) {
                    clb_spgemm_stats ss;
                    spg::spgemm_dispatch(ctl, mats[triples[3 * t + 1]], mats[triples[3 * t + 2]], so, prods[t], ss);
                    st.products += ss.products;
                    ++st.spgemm_calls;
                }
                for (u32 t = 0; t < ntriples; ++t) {
                    clb_dcsr &x = mats[triples[3 * t]];
                    clb_dcsr u{};
                    const clb_status rc = clb_dcsr_addition(ctl, &x, &prods[t], &u);
                    if (rc != CLB_SUCCESS) raise(static_cast<Status>(rc - 1), clb_last_error());
                    clb_dcsr_release(&x);
                    x = u;
                }
            } catch (...) {
                for (auto &p : prods) clb_dcsr_release(&p);
                throw;
            }
            for (auto &p : prods) clb_dcsr_release(&p);
        } else {
            if ((u64)ntriples * n > 0xFFFFFFFFull || (u64)xs.size() * n > 0xFFFFFFFFull)
                raise(INVALID_VALUE, "stacked operands exceed uint32 indices");
            // Y' (row block t = Y_t, columns shifted by t*n) and Z' (row block t = Z_t)
            std::vector<Seg> ys, zs;
            for (u32 t = 0; t < ntriples; ++t) {
                ys.push_back(seg_of(mats[triples[3 * t + 1]], t * n, t * n));
                zs.push_back(seg_of(mats[triples[3 * t + 2]], t * n, 0));
            }
            clb_dcsr Y = stack(ctl, ys, ntriples * n, ntriples * n);
            clb_dcsr Z{}, T{};
            try {
                Z = stack(ctl, zs, ntriples * n, n);
                clb_spgemm_stats ss;
                spg::spgemm_dispatch(ctl, Y, Z, so, T, ss);
                st.products = ss.products;
                st.spgemm_calls = 1;
            } catch (...) {
                clb_dcsr_release(&Y);
                clb_dcsr_release(&Z);
                throw;
            }
            clb_dcsr_release(&Y);
            clb_dcsr_release(&Z);
            // keys of every X block (rows shifted by xid*n) and of every product
            // block (its rows t*n + i re-based onto its X's block)
            u64 nkeys = T.nnz;
            for (u32 x : xs) nkeys += mats[x].nnz;
            DevArray<u64> keys(ctl, nkeys ? nkeys : 1);
            u64 pos = 0;
            try {
                for (u32 k = 0; k < (u32)xs.size(); ++k) {
                    const clb_dcsr &x = mats[xs[k]];
                    if (x.nnz)
                        k_keys_shifted<<<div_up(x.nzr, 4), 256, 0, ctl->stream>>>(u32p(x.rpt), u32p(x.rows),
                                                                                   u32p(x.cols), x.nzr, k * n,
                                                                                   keys.ptr + pos);
                    pos += x.nnz;
                }
                if (T.nnz) {
                    // T's rows are t*n + i; re-base each product block onto its X block
                    DevArray<u64> tk(ctl, T.nnz);
                    k_keys_shifted<<<div_up(T.nzr, 4), 256, 0, ctl->stream>>>(u32p(T.rpt), u32p(T.rows),
                                                                               u32p(T.cols), T.nzr, 0, tk.ptr);
                    CLB_LAUNCHED("keys");
                    // rows of T are sorted; block t occupies keys with row in [t*n, (t+1)*n)
                    DevArray<u64> tb(ctl, ntriples + 1);
                    k_block_bounds<<<div_up(ntriples + 1, 256), 256, 0, ctl->stream>>>(tk.ptr, T.nnz, n, ntriples,
                                                                                        tb.ptr);
                    std::vector<u64> hb(ntriples + 1);
                    readback(ctl, hb.data(), tb.ptr, 8 * (ntriples + 1));
                    for (u32 t = 0; t < ntriples; ++t) {
                        const u64 cnt = hb[t + 1] - hb[t];
                        if (!cnt) continue;
                        // keys (t*n + i) << 32 | c  ->  (xid*n + i) << 32 | c
                        const i64 delta = ((i64)xid(triples[3 * t]) - (i64)t) * (i64)n;
                        rebase_keys(ctl, tk.ptr + hb[t], cnt, delta, keys.ptr + pos);
                        pos += cnt;
                    }
                }
            } catch (...) {
                clb_dcsr_release(&T);
                throw;
            }
            clb_dcsr_release(&T);
            // one sort + deduplication of all blocks, then split by X block
            sort_pair_keys(ctl, keys.ptr, pos, (u32)(xs.size() * n - 1), n - 1);
            clb_buffer_s *ur = nullptr, *uc = nullptr;
            DevArray<u64> uk;
            const u32 m = unique_keys(ctl, keys.ptr, pos, uk);
            DevArray<u64> bnd(ctl, xs.size() + 1);
            k_block_bounds<<<div_up(xs.size() + 1, 256), 256, 0, ctl->stream>>>(uk.ptr, m, n, (u32)xs.size(),
                                                                                  bnd.ptr);
            CLB_LAUNCHED("block_bounds");
            std::vector<u64> hb(xs.size() + 1);
            readback(ctl, hb.data(), bnd.ptr, 8 * hb.size());
            (void)ur;
            (void)uc;
            for (u32 k = 0; k < (u32)xs.size(); ++k) {
                const u64 cnt = hb[k + 1] - hb[k];
                clb_dcsr &x = mats[xs[k]];
                if (cnt == x.nnz) continue;  // unions only add: equal size means unchanged
                clb_dcsr nx{};
                nx.nrows = n;
                nx.ncols = n;
                DevArray<u32> rr(ctl, cnt);
                nx.cols = buffer_new(ctl, 4 * cnt);
                k_unstack_pairs<<<div_up(cnt, 256), 256, 0, ctl->stream>>>(uk.ptr, hb[k], (u32)cnt, k * n, rr.ptr,
                                                                            u32p(nx.cols));
                CLB_LAUNCHED("unstack");
                rows_to_dcsr(ctl, rr.ptr, (u32)cnt, &nx.rpt, &nx.rows, &nx.nzr);
                nx.nnz = (u32)cnt;
                clb_dcsr_release(&x);
                x = nx;
            }
            sync(ctl);
        }
        for (u32 k = 0; k < nmats; ++k)
            if (mats[k].nnz != before[k]) {
                ++st.changed;
                st.added += mats[k].nnz - before[k];
            }
        st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
    });
}

}  // extern "C"
