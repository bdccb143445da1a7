// plan.hip -- a whole product A * B taken in row blocks, planned by the
// library from what each SpGEMM mode knows.
//
// A product larger than one uint32 DCSR (clBool's matrix_dcsr holds at most
// 2^32 - 1 nonzeros; its hash path silently wraps, hash.cpp:312,318) is taken
// as consecutive blocks of A's stored rows.  Where the block boundaries come
// from is part of each mode's cost, so the planning pass runs over all rows
// of A once, inside the same call sequence as the blocks:
//  * plain: the count and the symbolic pass over every row (exact row sizes;
//    the blocks then run only the numeric pass), cut by exact sizes;
//  * learned: features and predicted capacities of every row, cut by
//    capacities (a block whose exact result exceeds the bound is re-cut by the
//    exact sizes it found and run again);
//  * bound: min(products, ncols) per row, cut by that bound.
// The B row map and split table are built once and shared by all blocks.
// No counterpart in the reference.
#include "spgemm.hpp"

#include <algorithm>
#include <chrono>

struct clb_spgemm_plan_s {
    clb_controls_s *ctl = nullptr;
    clb_dcsr a{}, b{};
    clb::u32 mode = 0;
    clb::u64 bound = 0;
    bool empty = false;
    clb::spg::Setup s;                // over all stored rows of A
    clb::DevArray<clb::u32> size;     // exact sizes (plain) or capacities (one pass)
    clb::DevArray<float> pred;        // learned: predicted sizes
    std::vector<clb::u64> ppre, spre;  // host prefixes of products and sizes
    std::vector<clb::u32> cuts;       // block k = stored rows [cuts[k], cuts[k+1])
    clb_spgemm_plan_info info{};
};

namespace clb {
namespace spg {

static void retain_dcsr(const clb_dcsr &m) {
    if (m.rpt) buffer_retain(m.rpt);
    if (m.rows) buffer_retain(m.rows);
    if (m.cols) buffer_retain(m.cols);
}

// Greedy cut of rows [r0, r1) so that every block's size sum is at most lim
// (a row larger than lim is a block of its own).  pre[i] - pre[r0] is the
// size of rows [r0, i).  Appends the ends of the blocks to cuts.
static void cut_rows(const u64 *pre, u32 r0, u32 r1, u64 lim, std::vector<u32> &cuts) {
    u32 i = r0;
    while (i < r1) {
        const u64 *it = std::upper_bound(pre + i + 1, pre + r1 + 1, pre[i] + lim);
        u32 j = static_cast<u32>(it - pre) - 1;
        if (j <= i) j = i + 1;
        cuts.push_back(j);
        i = j;
    }
}

static void plan_create(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 mode, u64 block_bound,
                        clb_spgemm_plan_s &p) {
    check_dims(a, b);
    if (mode != CLB_SPGEMM_PLAIN && mode != CLB_SPGEMM_LEARNED && mode != CLB_SPGEMM_BOUND)
        raise(INVALID_ARGUMENT, "unknown SpGEMM mode");
    const auto t0 = std::chrono::steady_clock::now();
    p.ctl = ctl;
    p.mode = mode;
    p.bound = block_bound && block_bound < 0xFFFFFFFFull ? block_bound : 0xFFFFFFFFull;
    p.a = a;
    p.b = b;
    const u32 n = a.nzr;
    p.cuts = {0};
    p.empty = a.nnz == 0 || b.nnz == 0 || n == 0;
    if (p.empty) {
        // one block: its run returns the empty result (hash.cpp:64-67)
        p.cuts.push_back(n);
    } else {
        p.size.reset(ctl, n);
        if (mode == CLB_SPGEMM_PLAIN) {
            prepare(ctl, a, b, 0, n, p.s, true);
            plan_sizes_plain(ctl, p.s, p.size.ptr);
        } else {
            if (mode == CLB_SPGEMM_LEARNED) p.pred.reset(ctl, n);
            p.s.nrows = n;
            p.s.row0 = 0;
            plan_sizes_learned(ctl, a, b, p.s, p.size.ptr, p.pred.ptr, mode == CLB_SPGEMM_BOUND);
        }
        // the split table of B for every block (built here once when B's
        // columns span several bitmap windows)
        BinHost all{};
        all.off[NBINS] = 1;
        ensure_splits(ctl, p.s, all);
        DevArray<u64> pp(ctl, (u64)n + 1), sp(ctl, (u64)n + 1);
        exclusive_scan_u32_to_u64(ctl, p.s.P, pp.ptr, n, pp.ptr + n);
        exclusive_scan_u32_to_u64(ctl, p.size.ptr, sp.ptr, n, sp.ptr + n);
        p.ppre.resize((size_t)n + 1);
        p.spre.resize((size_t)n + 1);
        CLB_HIP(hipMemcpyAsync(p.ppre.data(), pp.ptr, 8ull * (n + 1), hipMemcpyDeviceToHost, ctl->stream),
                READ_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(p.spre.data(), sp.ptr, 8ull * (n + 1), hipMemcpyDeviceToHost, ctl->stream),
                READ_BUFFER_ERROR);
        sync(ctl);
        cut_rows(p.spre.data(), 0, n, p.bound, p.cuts);
        p.info.products = p.ppre[n];
        p.info.planned_nnz = p.spre[n];
    }
    // the plan keeps the operands alive
    retain_dcsr(a);
    retain_dcsr(b);
    p.info.mode = mode;
    p.info.nblocks = static_cast<u32>(p.cuts.size() - 1);
    p.info.ms_plan = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Block k exceeded the bound: size[cuts[k], cuts[k+1]) holds its exact row
// sizes now; cut those rows again.
static void recut(clb_controls_s *ctl, clb_spgemm_plan_s &p, u32 k) {
    const u32 rb = p.cuts[k], re = p.cuts[k + 1];
    std::vector<u32> sz(re - rb);
    readback(ctl, sz.data(), p.size.ptr + rb, 4ull * (re - rb));
    std::vector<u64> pre(re - rb + 1, 0);
    for (u32 i = 0; i < re - rb; ++i) pre[i + 1] = pre[i] + sz[i];
    std::vector<u32> sub;
    cut_rows(pre.data(), 0, re - rb, p.bound, sub);
    std::vector<u32> ends;
    for (u32 e : sub) ends.push_back(rb + e);
    ends.pop_back();  // == re, already cuts[k + 1]
    p.cuts.insert(p.cuts.begin() + k + 1, ends.begin(), ends.end());
    p.info.replans += 1;
    p.info.nblocks = static_cast<u32>(p.cuts.size() - 1);
}

static PlanRows rows_of(clb_spgemm_plan_s &p, u32 k, bool *replan) {
    const u32 rb = p.cuts[k], re = p.cuts[k + 1];
    return PlanRows{&p.s, p.s.P + rb, p.size.ptr + rb, p.pred.ptr ? p.pred.ptr + rb : nullptr,
                    p.ppre[re] - p.ppre[rb],
                    // a single row is never cut further
                    re - rb > 1 ? p.bound : 0xFFFFFFFFull, replan};
}

static void check_run(clb_controls_s *ctl, const clb_spgemm_plan_s &p, u32 k) {
    if (ctl != p.ctl) raise(INVALID_ARGUMENT, "a plan runs on the controls it was created with");
    if (k + 1 >= p.cuts.size()) raise(INVALID_ARGUMENT, "block index outside the plan");
}

static void plan_run(clb_controls_s *ctl, clb_spgemm_plan_s &p, u32 k, bool timing, clb_dcsr &out,
                     clb_spgemm_stats &st) {
    check_run(ctl, p, k);
    std::memset(&st, 0, sizeof(st));
    if (p.empty) {
        out = clb_dcsr{};
        out.nrows = p.a.nrows;
        out.ncols = p.b.ncols;
        ctl->last = st;
        return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const u32 rb = p.cuts[k], re = p.cuts[k + 1];
        bool replan = false;
        PlanRows pr = rows_of(p, k, &replan);
        clb_dcsr r{};
        std::memset(&st, 0, sizeof(st));
        if (p.mode == CLB_SPGEMM_PLAIN)
            spgemm_plain(ctl, p.a, p.b, rb, re, timing, r, st, &pr);
        else
            spgemm_learned(ctl, p.a, p.b, rb, re, timing, r, st, p.mode == CLB_SPGEMM_BOUND, &pr);
        if (!replan) {
            out = r;
            break;
        }
        recut(ctl, p, k);
    }
    if (!timing)
        st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctl->last = st;
}

// Every block in order, each handed to fn as soon as it is complete.  In
// the one-pass modes block k + 1's passes start while block k's tail
// (compaction and dense-row emission) still runs on the tail streams, and
// block k + 1's row kernels are launched before the host reads block k's
// results back, so that they follow block k's on the device without a gap.
static void plan_run_all(clb_controls_s *ctl, clb_spgemm_plan_s &p, bool timing, clb_spgemm_block_fn fn,
                         void *user) {
    if (ctl != p.ctl) raise(INVALID_ARGUMENT, "a plan runs on the controls it was created with");
    bool go = true;
    auto deliver = [&](u32 k, clb_dcsr &blk, clb_spgemm_stats &st) {
        ctl->last = st;
        if (!go) {  // stopped: results started before the stop are dropped
            clb_dcsr_release(&blk);
            return;
        }
        go = fn(user, k, &blk, &st) == 0;  // fn owns blk now
    };
    if (p.empty || p.mode == CLB_SPGEMM_PLAIN) {
        for (u32 k = 0; go && k + 1 < p.cuts.size(); ++k) {
            clb_dcsr r{};
            clb_spgemm_stats st;
            plan_run(ctl, p, k, timing, r, st);
            deliver(k, r, st);
        }
        return;
    }
    OnePassPtr pend, cur, nxt;  // pend: posted, its tail running; cur, nxt: launched
    u32 pend_k = 0;
    auto finish_pending = [&]() {
        clb_dcsr r{};
        clb_spgemm_stats st;
        one_pass_finish(*pend, r, st);
        pend.reset();
        deliver(pend_k, r, st);
    };
    auto avail = [&]() -> u64 {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) (void)hipGetLastError();
        return free_b + pool_for(ctl->device).cached_bytes();
    };
    auto launch = [&](u32 k) {
        const PlanRows pr = rows_of(p, k, nullptr);
        return one_pass_launch(ctl, p.a, p.b, p.cuts[k], p.cuts[k + 1], timing, p.mode == CLB_SPGEMM_BOUND, &pr);
    };
    u32 k = 0;
    while (go && k + 1 < p.cuts.size()) {
        if (!cur) {
            // overlap only when there is room for a second block of about
            // the pending one's size beside it (with room to spare for its
            // dense-row dump); otherwise let the pending block finish first
            // and hand its memory back to the pool
            if (pend && avail() < 3 * one_pass_bytes(*pend)) finish_pending();
            cur = launch(k);
        }
        // launch block k + 1 before posting block k, so that its row kernels
        // queue behind block k's on the side streams and the device does not
        // wait for the host between the blocks: when free memory holds twice
        // the larger of the pending block and block k with its output columns
        // (one such block for k + 1, and as much again for block k's output
        // and k + 1's later allocations)
        if (go && k + 2 < p.cuts.size()) {
            const PlanRows pr = rows_of(p, k, nullptr);
            const u64 mine = one_pass_bytes(*cur) + 4 * std::min<u64>(pr.products, pr.bound);
            const u64 need = std::max<u64>(pend ? one_pass_bytes(*pend) : 0, mine);
            if (avail() >= 2 * need) nxt = launch(k + 1);
        }
        one_pass_post(*cur);
        if (one_pass_replan(*cur)) {
            // cut block k finer and start again from it (block k + 1's
            // launch is dropped: its index moves with the new cuts)
            nxt.reset();
            cur.reset();
            recut(ctl, p, k);
            continue;
        }
        if (pend) finish_pending();
        pend = std::move(cur);
        pend_k = k;
        cur = std::move(nxt);
        ++k;
    }
    nxt.reset();
    cur.reset();  // launched after a stop: dropped
    if (pend) finish_pending();
}

}  // namespace spg
}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_spgemm_plan_create(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, uint32_t mode,
                                  uint64_t block_bound, clb_spgemm_plan *out) {
    return run_guarded([&] {
        if (!ctl || !a || !b || !out) raise(INVALID_ARGUMENT, "null argument");
        auto *p = new clb_spgemm_plan_s();
        try {
            spg::plan_create(ctl, *a, *b, mode, block_bound, *p);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

clb_status clb_spgemm_plan_info_get(clb_spgemm_plan plan, clb_spgemm_plan_info *info) {
    return run_guarded([&] {
        if (!plan || !info) raise(INVALID_ARGUMENT, "null argument");
        *info = plan->info;
    });
}

clb_status clb_spgemm_plan_block(clb_spgemm_plan plan, uint32_t k, uint32_t *row_begin, uint32_t *row_end) {
    return run_guarded([&] {
        if (!plan) raise(INVALID_ARGUMENT, "null plan");
        if (k + 1 >= plan->cuts.size()) raise(INVALID_ARGUMENT, "block index outside the plan");
        if (row_begin) *row_begin = plan->cuts[k];
        if (row_end) *row_end = plan->cuts[k + 1];
    });
}

clb_status clb_spgemm_plan_run(clb_controls ctl, clb_spgemm_plan plan, uint32_t k, uint32_t timing, clb_dcsr *out,
                               clb_spgemm_stats *stats) {
    return run_guarded([&] {
        if (!plan || !out) raise(INVALID_ARGUMENT, "null argument");
        clb_spgemm_stats st;
        clb_dcsr r{};
        spg::plan_run(ctl, *plan, k, timing != 0, r, st);
        *out = r;
        if (stats) *stats = st;
    });
}

clb_status clb_spgemm_plan_run_all(clb_controls ctl, clb_spgemm_plan plan, uint32_t timing, clb_spgemm_block_fn fn,
                                   void *user) {
    return run_guarded([&] {
        if (!plan || !fn) raise(INVALID_ARGUMENT, "null argument");
        spg::plan_run_all(ctl, *plan, timing != 0, fn, user);
    });
}

clb_status clb_spgemm_plan_release(clb_spgemm_plan plan) {
    return run_guarded([&] {
        if (!plan) return;
        (void)hipStreamSynchronize(plan->ctl->stream);
        clb_dcsr_release(&plan->a);
        clb_dcsr_release(&plan->b);
        delete plan;
    });
}

}  // extern "C"
