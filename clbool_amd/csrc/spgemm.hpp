// spgemm.hpp -- host-side pieces of the SpGEMM shared by the plain, learned
// and batched drivers.
#pragma once

#include "spgemm_kernels.hpp"

namespace clb {
namespace spg {

// Per-call device state: the B row map, per-row product counts and a few
// device scalars ([0] products, [1] nnz, [2] nzr, [3] overflow flag, ...).
struct Setup {
    u32 nrows = 0, row0 = 0, mapn = 0;
    DevArray<uint2> brange;
    DevArray<u32> bpres;  // presence bits of B rows (hypersparse B only)
    DevArray<u32> prod;
    DevArray<u64> scal;
    DevArray<u32> ctr;  // row counters of persistent launches (zeroed per launch)
    DevArray<u32> splits;  // per-slice B row ranges (Mats::splits) when ncols spans several windows
    Mats g{};
    BinCfg cfg{};
    clb_dcsr b{};  // the B operand (its buffers stay owned by the caller)
    u32 *P = nullptr;  // products per row of the block: prod.ptr, or a plan's array
};

// What a plan (plan.hip) hands to one row block's run: the block's slices of
// the per-row arrays computed over all of A at planning time, and the B row
// map and split table built once for all blocks.
struct PlanRows {
    const Setup *shared;  // B row map, presence bits, split table, bin config
    u32 *prod;            // products of the block's rows
    u32 *size;            // plain: exact row sizes; learned / bound: capacities
    const float *pred;    // learned: predicted sizes (nullptr otherwise)
    u64 products;         // sum of prod over the block
    u64 bound;            // one pass: a result above this many nonzeros is not
                          // assembled; the exact row sizes are written to
                          // `size` and *replan is set instead
    bool *replan;
};

// HIP-event timing of individual kernel launches, grouped by family
// (CLB_KF(pass, kind) of clbool_c.h).  Disabled unless the caller asked for timing.
struct LaunchTimer {
    bool on = false;
    struct Rec { int family; hipEvent_t a, b; };
    std::vector<Rec> recs;
    hipEvent_t pending = nullptr;
    void begin(hipStream_t st);
    void end(int family, hipStream_t st);
    void collect(clb_spgemm_stats &st);  // after the streams are synchronised
    ~LaunchTimer();
};

// Device layout of the binning scratch: hist[NBINS] | cursor[NBINS+1] | pad |
// products[NBINS] | a_nnz[NBINS] | out_nnz[NBINS] (the last three as u64).
struct BinBuf {
    DevArray<u32> buf;
    u32 *hist() const { return buf.ptr; }
    u32 *cursor() const { return buf.ptr + NBINS; }
    u64 *sums() const { return reinterpret_cast<u64 *>(buf.ptr + 2 * NBINS + 2); }
};
// Host copy of the binning: off[b]..off[b+1] is bin b's range of perm.
struct BinHost {
    u32 off[NBINS + 1];
    u64 products[NBINS], a_nnz[NBINS], out_nnz[NBINS];
};

// Scratch of the dense-bitmap path (see Dump in spgemm_kernels.hpp).
struct DumpBufs {
    bool on = false;
    u32 nbm = 0, nch = 0, words = 0, slots = 0, slice_words = 0;
    DevArray<uint4> store;
    DevArray<u32> cnt, list, count, isdense, fmt;
};
// Allocates the dense-bitmap scratch for the bitmap bin of `h` (nothing when
// the window is below 2^14 words, the knob disables it or it would not fit)
// and fills o.dump.  With perm and key (the rows' capacities or exact
// sizes), slots only for the rows likely to be dumped; else one per bitmap row.
void setup_dump(clb_controls_s *ctl, const Setup &s, const BinHost &h, DumpBufs &d, Out &o, bool one_pass,
                const u32 *perm = nullptr, const u32 *key = nullptr);
// Expands the dumped rows into cols at the exact row offsets `off`.
void emit_dense(clb_controls_s *ctl, const u32 *perm, const BinHost &h, DumpBufs &d, const u64 *off, u32 *cols,
                LaunchTimer *lt, int pass, clb_spgemm_stats *st);
// The same in two steps: the count and the chunk prefix into gpre on the
// controls' stream (returns the number of dumped rows), then the emission
// kernel on stream `es` (which must be ordered after the prefix).
u32 emit_prepare(clb_controls_s *ctl, DumpBufs &d, DevArray<u64> &gpre, clb_spgemm_stats *st);
void emit_launch(hipStream_t es, u32 count, const u32 *perm, const BinHost &h, const DumpBufs &d, const u64 *gpre,
                 const u64 *off, u32 *cols, LaunchTimer *lt, int pass);

BinCfg make_bincfg(const clb_controls_s *ctl, u32 ncols);
void ensure_splits(clb_controls_s *ctl, Setup &s, const BinHost &h);
// B row map, kernel arguments and (count == true) per-row product counts.
// With `shared`, the B row map, split table and bin config are taken from
// it (a plan's Setup) instead of being built.
void prepare(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 nrows, Setup &s,
             bool count = true, const Setup *shared = nullptr);
// Bins the rows of the block by (P, key) into perm (device only, no sync).
void bin_rows(clb_controls_s *ctl, const Setup &s, const u32 *key, u32 *nnz_row_zero, DevArray<u32> &perm,
              BinBuf &bb);
// Reads the binning back (synchronises the stream).
void read_bins(clb_controls_s *ctl, const BinBuf &bb, BinHost &h);
// Per-bin sums of nnz_row over the rows of each nonempty bin's range of perm
// (rows marked FAILED count 0); the result is read back into out[NBINS].
void sum_by_bin(clb_controls_s *ctl, const u32 *perm, const BinHost &h, u32 nrows, const u32 *nnz_row,
                u64 out[NBINS]);
// Launches the row kernels of every nonempty bin in `mode` (SYM, NUM or ONE)
// on the side streams.  family < 0 times launches as sym_* / num_* by mode.
// The controls' stream waits for them at the end, unless `defer` is given:
// then the end of each side stream's work is recorded in defer[k] (NSIDE
// events, made by the caller) for join_deferred.  one_stream: every kernel
// on the controls' stream instead (no side streams).
void launch_bins(clb_controls_s *ctl, const Setup &s, const u32 *perm, const u32 off[NBINS + 1], int mode,
                 const Out &o, LaunchTimer *lt, int family, hipEvent_t *defer = nullptr, bool one_stream = false);
// The controls' stream waits for the work recorded by a deferred launch_bins.
void join_deferred(clb_controls_s *ctl, const hipEvent_t *defer);
void scan_sizes(clb_controls_s *ctl, Setup &s, const u32 *nnz_row, DevArray<u64> &off, DevArray<u32> &pos);
void assemble(clb_controls_s *ctl, const Setup &s, const clb_dcsr &a, const u32 *nnz_row, const u64 *off,
              const u32 *pos, u64 nnz, u32 nzr, clb_buffer_s *cols, clb_dcsr &out);
void check_dims(const clb_dcsr &a, const clb_dcsr &b);

// pr != nullptr: the block of a plan (the count and symbolic pass were run
// over all rows at planning time; pr->size holds the exact row sizes).
void spgemm_plain(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                  bool timing, clb_dcsr &out, clb_spgemm_stats &st, const PlanRows *pr = nullptr);
// Planning passes over all stored rows of A (s from prepare over all rows):
// exact row sizes by the symbolic pass (plain) into size.
void plan_sizes_plain(clb_controls_s *ctl, Setup &s, u32 *size);
// learned.hip.  bound_only: capacities min(P, ncols) for every row, no
// predictor (CLB_SPGEMM_BOUND).
void spgemm_learned(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                    bool timing, clb_dcsr &out, clb_spgemm_stats &st, bool bound_only = false,
                    const PlanRows *pr = nullptr);
// The one pass of one block in two halves: start runs everything up to the
// tail (compaction and dense-row emission), which it leaves running on the
// controls' tail streams; finish waits for it and hands over the result.  A
// block that exceeded a plan's bound (one_pass_replan) has no result.
// start is itself launch (binning, and the row kernels left running on the
// side streams) followed by post (the rest, from the overflow count on), so
// that a caller can launch the next block before posting this one: its row
// kernels then queue behind this block's on the side streams and start as
// soon as they end, while the host posts this block.
struct OnePass;
struct OnePassDeleter {
    void operator()(OnePass *p) const;  // waits for the block's kernels first
};
using OnePassPtr = std::unique_ptr<OnePass, OnePassDeleter>;
OnePassPtr one_pass_start(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                          bool timing, bool bound_only, const PlanRows *pr);
OnePassPtr one_pass_launch(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                           bool timing, bool bound_only, const PlanRows *pr);
void one_pass_post(OnePass &p);
bool one_pass_replan(const OnePass &p);
u64 one_pass_bytes(const OnePass &p);  // device memory the block holds
void one_pass_finish(OnePass &p, clb_dcsr &out, clb_spgemm_stats &st);
// Planning pass of the one pass: s prepared over all rows (products in s.P
// afterwards), capacities into size, predictions into pred (learned only).
void plan_sizes_learned(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, Setup &s, u32 *size,
                        float *pred, bool bound_only);
// Copies every row r with nnz_row[r] != 0 to cols + off[r] (off: exact
// offsets, off[nrows] = nnz): from side + soff[r] when key2[r] != 0, else
// from temp + toff[r]; rows with isdense[r] are skipped (written elsewhere).
// With pred, accumulates |log2(1 + pred[r]) - log2(1 + nnz_row[r])| into
// *err.  by_tiles: workgroups take equal tiles of output columns instead of
// one row each.  (learned.hip)
void compact_rows(hipStream_t st, u32 nrows, u64 nnz, const u32 *nnz_row, const u64 *off, const u32 *temp,
                  const u64 *toff, const u32 *side, const u64 *soff, const u32 *key2, const u32 *isdense,
                  const float *pred, u32 *cols, double *err, bool by_tiles);
// esc.hip: the expand-sort-compress SpGEMM (dcsr::matrix_multiplication)
void spgemm_esc(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, clb_dcsr &out, clb_spgemm_stats &st);
void spgemm_dispatch(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, const clb_spgemm_opts &o,
                     clb_dcsr &out, clb_spgemm_stats &st);

}  // namespace spg
}  // namespace clb
