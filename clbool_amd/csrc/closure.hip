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
// rows of all products.  The product is taken in row blocks of Y' whose
// output bound (sum over rows of min(products, n)) stays below 2^32, so each
// block's result fits uint32 indices however large the step grows.  Each X's
// union merges sorted key segments -- its own keys and, per triple that
// targets it, that triple's product rows (already sorted: a product block is
// canonical DCSR) -- and drops duplicates.  Every step of both modes is
// bit-exact: the products
// and unions are exact set operations, only the launch structure differs.
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

// Keys (row << 32 | col) of stored rows [r0, r1) of a DCSR matrix, rows
// shifted (modulo 2^32), written from keys[rpt[r] - nnz_base]; one wave per row.
__global__ __launch_bounds__(256) void k_keys_range(const u32 *__restrict__ rpt, const u32 *__restrict__ rows,
                                                    const u32 *__restrict__ cols, u32 r0, u32 r1, u32 row_shift,
                                                    u32 nnz_base, u64 *__restrict__ keys) {
    const u32 r = r0 + blockIdx.x * 4 + dev::wave_id();
    if (r >= r1) return;
    const u64 hi = (u64)(rows[r] + row_shift) << 32;
    const u32 e = rpt[r + 1];
    for (u32 j = rpt[r] + dev::lane(); j < e; j += 64) keys[j - nnz_base] = hi | cols[j];
}

// zlen[row id] = stored length of that row of Z' (zlen pre-filled with 0).
__global__ void k_row_lengths(const u32 *__restrict__ rpt, const u32 *__restrict__ rows, u32 nzr,
                              u32 *__restrict__ zlen) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nzr) zlen[rows[r]] = rpt[r + 1] - rpt[r];
}

// bound[r] = min(products of stored row r of Y', n): an upper bound of the
// row's output size.  One wave per row (rows of Y' reach thousands of
// nonzeros).
__global__ __launch_bounds__(256) void k_row_bounds(const u32 *__restrict__ rpt, const u32 *__restrict__ cols,
                                                    u32 nzr, const u32 *__restrict__ zlen, u32 n,
                                                    u64 *__restrict__ bound) {
    const u32 r = blockIdx.x * 4 + dev::wave_id();
    if (r >= nzr) return;
    u64 p = 0;
    for (u32 j = rpt[r] + dev::lane(); j < rpt[r + 1]; j += dev::WAVE) p += zlen[cols[j]];
    p = dev::wave_sum(p);
    if (dev::lane() == 0) bound[r] = p < n ? p : n;
}

// cuts[b] = first stored row whose exclusive bound prefix is >= b * step
// (pre holds nzr + 1 entries, pre[nzr] = the total).  Every row's bound is at
// most n, so with step = lim - n + 1 a block [cuts[b], cuts[b+1]) has bound
// pre[cuts[b+1]] - pre[cuts[b]] < (b+1)*step + n - b*step, i.e. at most lim.
// bbound[b] receives that block bound.
__global__ void k_cuts(const u64 *__restrict__ pre, u32 nzr, u64 step, u32 nb, u32 *__restrict__ cuts) {
    const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    if (b == nb) {
        cuts[b] = nzr;
        return;
    }
    const u64 target = (u64)b * step;
    u32 lo = 0, hi = nzr;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo) / 2;
        if (pre[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    cuts[b] = lo;
}

__global__ void k_block_bounds(const u64 *__restrict__ pre, const u32 *__restrict__ cuts, u32 nb,
                               u64 *__restrict__ bbound) {
    const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) bbound[b] = pre[cuts[b + 1]] - pre[cuts[b]];
}

// For row blocks t = 0..nt of a product (row ids t*n + i): out[2t] = first
// stored row with id >= t*n, out[2t+1] = its position in cols.
__global__ void k_row_splits(const u32 *__restrict__ rpt, const u32 *__restrict__ rows, u32 nzr, u32 n, u32 nt,
                             u32 *__restrict__ out) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nt) return;
    const u64 target = (u64)t * n;
    u32 lo = 0, hi = nzr;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo) / 2;
        if ((u64)rows[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    out[2 * t] = lo;
    out[2 * t + 1] = rpt[lo];
}

// Output bound of one stacked row block: the caller's block_bound (0 or
// above 2^32 - 1: 2^32 - 1), raised to 2n when smaller (one row's bound is up
// to n, and cuts need room for at least one row).
static u64 closure_block_limit(u32 block_bound, u32 n) {
    u64 lim = block_bound ? block_bound : 0xFFFFFFFFull;
    if (lim < 2ull * n) lim = 2ull * n;
    return lim < 0xFFFFFFFFull ? lim : 0xFFFFFFFFull;
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
        auto t_prod = t0;  // end of the products phase

        if (o.mode == CLB_CLOSURE_PER_TRIPLE) {
            // the reference-style loop: products from the step's start values,
            // then one union per triple
            std::vector<clb_dcsr> prods(ntriples);
            try {
                for (u32 t = 0; t < ntriples; ++t) {
                    clb_spgemm_stats ss;
                    spg::spgemm_dispatch(ctl, mats[triples[3 * t + 1]], mats[triples[3 * t + 2]], so, prods[t], ss);
                    st.products += ss.products;
                    st.fallback_rows += ss.fallback_rows;
                    st.dense_rows += ss.dense_rows;
                    ++st.spgemm_calls;
                }
                sync(ctl);
                t_prod = std::chrono::steady_clock::now();
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
            if ((u64)ntriples * n > 0xFFFFFFFFull || n > (1u << 31))
                raise(INVALID_VALUE, "stacked operands exceed uint32 indices");
            // Y' (row block t = Y_t, columns shifted by t*n) and Z' (row block t = Z_t)
            std::vector<Seg> ys, zs;
            for (u32 t = 0; t < ntriples; ++t) {
                ys.push_back(seg_of(mats[triples[3 * t + 1]], t * n, t * n));
                zs.push_back(seg_of(mats[triples[3 * t + 2]], t * n, 0));
            }
            std::vector<clb_dcsr> prods;  // row blocks of Y' * Z'
            std::vector<std::vector<u32>> splits;
            auto release_all = [&] {
                for (auto &p : prods) clb_dcsr_release(&p);
            };
            clb_dcsr Y = stack(ctl, ys, ntriples * n, ntriples * n);
            clb_dcsr Z{};
            try {
                Z = stack(ctl, zs, ntriples * n, n);
                if (Y.nnz && Z.nnz) {
                    // row blocks of Y' with output bound below 2^32 each
                    DevArray<u32> zlen(ctl, (u64)ntriples * n);
                    fill_u32(ctl, zlen.ptr, (u64)ntriples * n, 0u);
                    k_row_lengths<<<div_up(Z.nzr, 256), 256, 0, ctl->stream>>>(u32p(Z.rpt), u32p(Z.rows), Z.nzr,
                                                                                zlen.ptr);
                    CLB_LAUNCHED("row_lengths");
                    DevArray<u64> pre(ctl, (u64)Y.nzr + 1);
                    k_row_bounds<<<div_up(Y.nzr, 4), 256, 0, ctl->stream>>>(u32p(Y.rpt), u32p(Y.cols), Y.nzr,
                                                                               zlen.ptr, n, pre.ptr);
                    CLB_LAUNCHED("row_bounds");
                    exclusive_scan_u64(ctl, pre.ptr, pre.ptr, Y.nzr, pre.ptr + Y.nzr);
                    u64 total = 0;
                    readback(ctl, &total, pre.ptr + Y.nzr, 8);
                    const u64 lim = closure_block_limit(o.block_bound, n);
                    const u64 step = lim - n + 1;
                    const u32 nb = (u32)((total + step - 1) / step);
                    DevArray<u32> cuts(ctl, (u64)nb + 1);
                    DevArray<u64> bb(ctl, nb ? nb : 1);
                    k_cuts<<<div_up(nb + 1, 256), 256, 0, ctl->stream>>>(pre.ptr, Y.nzr, step, nb, cuts.ptr);
                    CLB_LAUNCHED("cuts");
                    if (nb) {
                        k_block_bounds<<<div_up(nb, 256), 256, 0, ctl->stream>>>(pre.ptr, cuts.ptr, nb, bb.ptr);
                        CLB_LAUNCHED("block_bounds");
                    }
                    std::vector<u32> hc(nb + 1);
                    std::vector<u64> hb(nb);
                    readback(ctl, hc.data(), cuts.ptr, 4ull * (nb + 1));
                    if (nb) readback(ctl, hb.data(), bb.ptr, 8ull * nb);
                    for (u32 k = 0; k < nb; ++k) {
                        if (hb[k] > lim) raise(RUN_KERNEL_FAILURE, "closure row block exceeds its output bound");
                        if (hb[k] > st.max_block_bound) st.max_block_bound = hb[k];
                    }
                    DevArray<u32> sp(ctl, 2ull * (ntriples + 1));
                    auto absorb = [&](clb_dcsr &T, const clb_spgemm_stats &ss) {
                        st.products += ss.products;
                        st.fallback_rows += ss.fallback_rows;
                        st.dense_rows += ss.dense_rows;
                        ++st.spgemm_calls;
                        if (!T.nnz) {
                            clb_dcsr_release(&T);
                            return;
                        }
                        prods.push_back(T);
                        // where each product block t starts in T
                        k_row_splits<<<div_up(ntriples + 1, 256), 256, 0, ctl->stream>>>(
                            u32p(T.rpt), u32p(T.rows), T.nzr, n, ntriples, sp.ptr);
                        CLB_LAUNCHED("row_splits");
                        std::vector<u32> hs(2 * (ntriples + 1));
                        readback(ctl, hs.data(), sp.ptr, 4ull * hs.size());
                        splits.push_back(hs);
                    };
                    const bool one_pass = so.mode == CLB_SPGEMM_LEARNED || so.mode == CLB_SPGEMM_BOUND;
                    if (!one_pass) {
                        for (u32 k = 0; k < nb; ++k) {
                            if (hc[k + 1] <= hc[k]) continue;
                            clb_spgemm_opts bo = so;
                            bo.row_begin = hc[k];
                            bo.row_end = hc[k + 1];
                            clb_spgemm_stats ss;
                            clb_dcsr T{};
                            spg::spgemm_dispatch(ctl, Y, Z, bo, T, ss);
                            absorb(T, ss);
                        }
                    } else {
                        // one-pass blocks overlap: block k + 1's passes start while
                        // block k's compaction and dense-row emission still run,
                        // and its row kernels are launched before block k is
                        // posted, so that they follow block k's on the device
                        // (each when device memory allows)
                        spg::OnePassPtr pend, cur, nxt;
                        auto finish_pending = [&]() {
                            clb_dcsr T{};
                            clb_spgemm_stats ss;
                            spg::one_pass_finish(*pend, T, ss);
                            pend.reset();
                            absorb(T, ss);
                        };
                        auto avail = [&]() -> u64 {
                            size_t free_b = 0, total_b = 0;
                            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) (void)hipGetLastError();
                            return free_b + pool_for(ctl->device).cached_bytes();
                        };
                        auto launch = [&](u32 k) {
                            const bool bound_only = so.mode == CLB_SPGEMM_BOUND;
                            return spg::one_pass_launch(ctl, Y, Z, hc[k], hc[k + 1], false, bound_only, nullptr);
                        };
                        auto nonempty = [&](u32 k) {  // the first block from k with rows
                            while (k < nb && hc[k + 1] <= hc[k]) ++k;
                            return k;
                        };
                        for (u32 k = nonempty(0); k < nb;) {
                            if (!cur) {
                                if (pend && avail() < 3 * spg::one_pass_bytes(*pend)) finish_pending();
                                cur = launch(k);
                            }
                            const u32 k2 = nonempty(k + 1);
                            if (k2 < nb) {
                                const u64 mine = spg::one_pass_bytes(*cur) + 4 * hb[k];
                                const u64 need = std::max<u64>(pend ? spg::one_pass_bytes(*pend) : 0, mine);
                                if (avail() >= 2 * need) nxt = launch(k2);
                            }
                            spg::one_pass_post(*cur);
                            if (pend) finish_pending();
                            pend = std::move(cur);
                            cur = std::move(nxt);
                            k = k2;
                        }
                        if (pend) finish_pending();
                    }
                }
            } catch (...) {
                release_all();
                clb_dcsr_release(&Y);
                clb_dcsr_release(&Z);
                throw;
            }
            clb_dcsr_release(&Y);
            clb_dcsr_release(&Z);
            sync(ctl);
            t_prod = std::chrono::steady_clock::now();
            try {
                // one union per distinct X.  Its keys and, per triple that
                // targets it, the keys of that triple's product rows (its
                // parts in consecutive row blocks concatenate in row order)
                // form sorted segments; rounds of pairwise merges turn them
                // into one sorted list, whose duplicates are then dropped.
                for (u32 k = 0; k < (u32)xs.size(); ++k) {
                    clb_dcsr &x = mats[xs[k]];
                    u64 seg = x.nnz;
                    for (size_t q = 0; q < prods.size(); ++q)
                        for (u32 t = 0; t < ntriples; ++t)
                            if (triples[3 * t] == xs[k]) seg += splits[q][2 * t + 3] - splits[q][2 * t + 1];
                    if (seg == x.nnz) continue;
                    if (seg >= 0xFFFFFFFFull)
                        raise(INVALID_VALUE, "a closure union exceeds uint32 nonzeros");
                    DevArray<u64> keys(ctl, seg), tmp(ctl, seg);
                    std::vector<std::pair<u64, u64>> segs;  // (start, length) in keys
                    u64 pos = 0;
                    if (x.nnz) {
                        k_keys_shifted<<<div_up(x.nzr, 4), 256, 0, ctl->stream>>>(u32p(x.rpt), u32p(x.rows),
                                                                                   u32p(x.cols), x.nzr, 0, keys.ptr);
                        CLB_LAUNCHED("keys");
                        segs.push_back({0, x.nnz});
                        pos = x.nnz;
                    }
                    for (u32 t = 0; t < ntriples; ++t) {
                        if (triples[3 * t] != xs[k]) continue;
                        const u64 start = pos;
                        for (size_t q = 0; q < prods.size(); ++q) {
                            const clb_dcsr &T = prods[q];
                            const std::vector<u32> &hs = splits[q];
                            const u32 r0 = hs[2 * t], r1 = hs[2 * t + 2];
                            const u32 e0 = hs[2 * t + 1], e1 = hs[2 * t + 3];
                            if (e1 == e0) continue;
                            // rows t*n + i -> i
                            k_keys_range<<<div_up(r1 - r0, 4), 256, 0, ctl->stream>>>(
                                u32p(T.rpt), u32p(T.rows), u32p(T.cols), r0, r1, 0u - t * n,
                                (u32)(e0 - (u32)pos), keys.ptr);
                            CLB_LAUNCHED("keys_range");
                            pos += e1 - e0;
                        }
                        if (pos > start) segs.push_back({start, pos - start});
                    }
                    u64 *cur = keys.ptr, *nxt = tmp.ptr;
                    while (segs.size() > 1) {
                        std::vector<std::pair<u64, u64>> next;
                        for (size_t i = 0; i < segs.size(); i += 2) {
                            const auto &sa = segs[i];
                            if (i + 1 == segs.size()) {  // odd one out: copied to the other buffer
                                CLB_HIP(hipMemcpyAsync(nxt + sa.first, cur + sa.first, 8 * sa.second,
                                                       hipMemcpyDeviceToDevice, ctl->stream),
                                        COPY_BUFFER_ERROR);
                                next.push_back(sa);
                                continue;
                            }
                            const auto &sb = segs[i + 1];  // adjacent: sb.first == sa.first + sa.second
                            merge_keys(ctl, cur + sa.first, sa.second, cur + sb.first, sb.second, nxt + sa.first);
                            next.push_back({sa.first, sa.second + sb.second});
                        }
                        segs.swap(next);
                        std::swap(cur, nxt);
                    }
                    clb_dcsr nx{};
                    nx.nrows = n;
                    nx.ncols = n;
                    keys_to_dcsr(ctl, cur, pos, nx);
                    if (nx.nnz == x.nnz) {  // unions only add: equal size means unchanged
                        clb_dcsr_release(&nx);
                        continue;
                    }
                    clb_dcsr_release(&x);
                    x = nx;
                }
                sync(ctl);
            } catch (...) {
                release_all();
                throw;
            }
            release_all();
        }
        for (u32 k = 0; k < nmats; ++k)
            if (mats[k].nnz != before[k]) {
                ++st.changed;
                st.added += mats[k].nnz - before[k];
            }
        sync(ctl);
        const auto t1 = std::chrono::steady_clock::now();
        st.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st.ms_products = std::chrono::duration<double, std::milli>(t_prod - t0).count();
        st.ms_unions = std::chrono::duration<double, std::milli>(t1 - t_prod).count();
        if (stats) *stats = st;
    });
}

}  // extern "C"
