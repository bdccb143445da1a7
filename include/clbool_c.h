/*
 * clbool_c.h -- C-ABI of the MI355X-native boolean sparse-matrix library.
 *
 * This is the drop-in boundary for clBool's SpGEMM path.  Every entry point
 * names the reference interface it replaces (paths relative to the clBool
 * source tree).  The C++ header clbool.hpp wraps these calls into clBool's own
 * classes (clbool::Controls, matrix_dcsr, matrix_coo, matrix_csr, ...) so that
 * code written against include/clbool.hpp of the reference compiles unchanged
 * against this library.
 *
 * Conventions
 *  - Every call returns clb_status: 0 on success, otherwise (Status + 1) where
 *    Status is clBool's enum (src/core/error.hpp:130-151).  The message of the
 *    last failure on the calling thread is returned by clb_last_error().
 *  - Buffers are reference counted device allocations (they replace
 *    cl::Buffer).  Matrices are plain structs of buffer handles plus sizes;
 *    an output matrix is filled with NEW references that the caller owns and
 *    eventually drops with clb_*_release.  Inputs are borrowed.  An output may
 *    be the same struct as an input: the call reads all inputs first.
 *  - All operations are ordered on the Controls' stream and complete before
 *    the call returns, as in the reference (blocking reads and waits).
 *  - Indices and sizes are uint32_t (src/core/matrix_base.hpp:12-16).
 */
#ifndef CLBOOL_C_H
#define CLBOOL_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t clb_status;
#define CLB_SUCCESS 0

typedef struct clb_controls_s *clb_controls;
typedef struct clb_buffer_s *clb_buffer;

/* matrix_dcsr (src/core/matrix_dcsr.hpp:11-84) */
typedef struct clb_dcsr {
    clb_buffer rpt;   /* uint32[nzr + 1] */
    clb_buffer rows;  /* uint32[nzr], strictly ascending */
    clb_buffer cols;  /* uint32[nnz], ascending and unique within each row */
    uint32_t nrows, ncols, nnz, nzr;
} clb_dcsr;

/* matrix_coo (src/core/matrix_coo.hpp:11-77) */
typedef struct clb_coo {
    clb_buffer rows, cols; /* uint32[nnz] each */
    uint32_t nrows, ncols, nnz;
} clb_coo;

/* matrix_csr (src/core/matrix_csr.hpp:7-44) */
typedef struct clb_csr {
    clb_buffer rpt;   /* uint32[nrows + 1] */
    clb_buffer cols;  /* uint32[nnz] */
    uint32_t nrows, ncols, nnz;
} clb_csr;

/* Kernel-selection knobs of the SpGEMM (the hard-coded thresholds of
 * src/dcsr/dcsr_matrix_multiplication_hash.cpp:13-47, made tunable). */
typedef struct clb_tuning {
    uint32_t hash_max_slots;       /* largest LDS hash table (slots, pow2)        */
    uint32_t bitmap_min_products;  /* rows with at least this many products use
                                      the LDS bitmap accumulator                  */
    uint32_t bitmap_slot_ratio;    /* use a bitmap when its words <= ratio * the
                                      hash slots the row would need (when the
                                      columns fit one bitmap window)              */
    uint32_t tiny_max_products;    /* rows up to this many products use the
                                      wave-level sort kernel                      */
    uint32_t learned_margin_pct;   /* learned sizing: capacity = pred * (100 + m)%  */
    uint32_t emit_lane_bits_min;   /* bitmap emission: 64-word steps with at least
                                      this many set bits are written one lane per
                                      bit, sparser ones one lane per word
                                      (0: 1024)                                   */
    uint32_t dump_bits_per_chunk;  /* bitmap rows (windows of at least 2^14
                                      words) whose first window holds at least
                                      this many set bits per 1024 words are
                                      dumped and expanded by a separate emission
                                      kernel (>= 65536: off; 0: 1024 in the
                                      learned one pass, off in the numeric pass) */
    uint32_t bitmap_persistent;    /* nonzero: the bitmap rows run on a persistent
                                      grid that takes rows from a counter
                                      instead of one workgroup per row          */
    uint32_t reserved[8];          /* reserved[CLB_TUNE_BITMAP_PRED_CAPS]: nonzero:
                                      the one pass sizes bitmap rows by their
                                      predicted capacity too (a bitmap row over
                                      it is dumped whole, or re-run exactly),
                                      0: by min(products, ncols)
                                      reserved[CLB_TUNE_GATHER_COMBINE]: nonzero:
                                      the bitmap gather ORs the bits that lanes
                                      of a wave set in one 32-bit word together
                                      and issues one LDS atomic per word
                                      reserved[CLB_TUNE_ESC_MAX_PRODUCTS]: rows
                                      bound for a bitmap whose columns span
                                      several bitmap windows go to the LDS
                                      expand-sort-compress kernel when they have
                                      at most this many products (at most 8192;
                                      0: never)
                                      reserved[CLB_TUNE_COMPACT_TILES]: nonzero:
                                      the one pass's compaction gives every
                                      workgroup an equal tile of output columns
                                      (0: one workgroup per row)
                                      reserved[CLB_TUNE_HASH_WIDE]: bit b set:
                                      LDS hash tables of 2^(7+b) slots run with
                                      twice the default workgroup (64 threads
                                      for 2^7..2^9 slots, 128 for 2^10..2^11,
                                      256 for 2^12, 512 for 2^13; 2^14: 1024
                                      either way); bit
                                      CLB_TUNE_TABLES_OVER_SLICES_BIT set: when
                                      the columns span several bitmap windows,
                                      bitmap_slot_ratio sends no row to a bitmap
                                      (rows within hash_max_slots take a table)
                                      reserved[CLB_TUNE_ONEPASS_BITMAP_MIN]: in
                                      the one pass, rows with at least this many
                                      products use a bitmap whatever their
                                      predicted size (0: the size alone decides)
                                      reserved[CLB_TUNE_BITMAP_WINDOW_BITS]:
                                      13, 14 or 15: the numeric and one-pass
                                      bitmap kernels use windows of 2^this
                                      32-bit words (13 and 14: 512-thread
                                      workgroups, two or more per CU; 15: one
                                      1024-thread workgroup per CU), capped at
                                      the matrix width; 0: 2^14 up to 2^20
                                      columns, else 2^15
                                      reserved[CLB_TUNE_EMIT_SPARSE_BITS]: a
                                      bitmap window slice with fewer set bits
                                      than this per 1024 words is written by
                                      each lane from its own words (one wave
                                      scan per 256 words, empty ones skipped);
                                      denser slices 64 words per step (0: 256) */
} clb_tuning;
#define CLB_TUNE_BITMAP_PRED_CAPS 0
#define CLB_TUNE_GATHER_COMBINE   1
#define CLB_TUNE_ESC_MAX_PRODUCTS 2
#define CLB_TUNE_COMPACT_TILES    3
#define CLB_TUNE_HASH_WIDE        4
#define CLB_TUNE_TABLES_OVER_SLICES_BIT 16  /* a bit of reserved[CLB_TUNE_HASH_WIDE] */
#define CLB_TUNE_ONEPASS_BITMAP_MIN 5
#define CLB_TUNE_BITMAP_WINDOW_BITS 6
#define CLB_TUNE_EMIT_SPARSE_BITS   7

#define CLB_SPGEMM_PLAIN   0u  /* two-pass: symbolic count, scan, numeric      */
#define CLB_SPGEMM_LEARNED 1u  /* one pass sized by the learned predictor,
                                  overflowing rows re-run exactly              */
#define CLB_SPGEMM_BOUND   2u  /* one pass with every row sized by its upper
                                  bound min(products, ncols): no predictor,
                                  never a re-run (the learned mode without
                                  its network)                                 */

typedef struct clb_spgemm_opts {
    uint32_t mode;       /* CLB_SPGEMM_PLAIN, _LEARNED or _BOUND               */
    uint32_t row_begin;  /* range of A's stored rows to multiply (row block);  */
    uint32_t row_end;    /* row_end == 0 means all rows                       */
    uint32_t timing;     /* nonzero: record per-stage times (HIP events)      */
} clb_spgemm_opts;

#define CLB_MAX_BINS 32
typedef struct clb_spgemm_stats {
    uint64_t products;      /* sum over rows of sum_{k in A_i} |B_k|          */
    uint64_t nnz;           /* output nonzeros                               */
    uint32_t nzr;           /* output nonempty rows                          */
    uint32_t rows;          /* A rows processed                              */
    uint32_t fallback_rows; /* learned mode: rows whose table overflowed     */
    uint32_t n_bins;
    uint32_t bin_rows[CLB_MAX_BINS];     /* rows per bin (numeric binning)          */
    uint64_t bin_products[CLB_MAX_BINS]; /* intermediate products per bin           */
    uint64_t bin_a_nnz[CLB_MAX_BINS];    /* A nonzeros of the bin's rows            */
    uint64_t bin_out_nnz[CLB_MAX_BINS];  /* output nonzeros of the bin's rows       */
    double   pred_abs_log2_err; /* learned mode: mean |log2(1+pred)-log2(1+nnz)| */
    double   ms_setup, ms_symbolic, ms_numeric, ms_assemble, ms_total;
    /* with opts.timing: HIP-event durations summed per kernel family, indexed
     * by CLB_KF(pass, kind); launches[] counts the launches of each family */
    double   ms_kernel[16];
    uint32_t launches[16];
    uint64_t capacity;      /* learned mode: columns reserved by the predicted sizes */
    uint64_t dense_rows;    /* bitmap rows written by the dense-row emission kernel */
    uint64_t dense_nnz;     /* their output nonzeros                         */
} clb_spgemm_stats;

/* kernel families of clb_spgemm_stats.ms_kernel: pass x kind */
#define CLB_KP_SYMBOLIC 0u  /* plain mode, symbolic pass               */
#define CLB_KP_NUMERIC  1u  /* plain mode, numeric pass                */
#define CLB_KP_ONEPASS  2u  /* learned / bound mode, the single sized pass */
#define CLB_KP_FALLBACK 3u  /* learned mode, exact re-run of overflows */
#define CLB_KK_TINY     0u
#define CLB_KK_HASH     1u  /* LDS hash tables and LDS expand-sort-compress rows */
#define CLB_KK_BITMAP   2u
#define CLB_KK_EMIT     3u  /* emission of dumped dense bitmap rows    */
#define CLB_KF(pass, kind) ((pass) * 4u + (kind))
/* The re-run never dumps rows (it has no emission launch), so its EMIT slot
 * holds the one pass's compaction instead: every row copied from its
 * temporary slot to its exact position. */
#define CLB_KF_COMPACT CLB_KF(CLB_KP_FALLBACK, CLB_KK_EMIT)

typedef struct clb_device_info {
    char     name[256];
    char     arch[64];
    uint32_t compute_units;
    uint32_t max_wg_size;
    uint32_t lds_bytes;
    uint64_t global_mem_bytes;
    uint32_t clock_khz;
    uint32_t mem_clock_khz;
    uint32_t mem_bus_width;
} clb_device_info;

/* ---- errors ------------------------------------------------------------ */
/* Message of the last failure on this thread (Exception::what(),
 * src/core/error.hpp:200-208). */
const char *clb_last_error(void);

/* ---- runtime (replaces src/core/controls.hpp, src/common/env.cpp) ------ */
clb_status clb_device_count(uint32_t *count);
/* create_controls (src/common/env.cpp:7-36): platform 0 is the HIP runtime;
 * an unknown platform or device id fails with CLBOOL_INITIALIZATION_ERROR. */
clb_status clb_controls_create(uint32_t platform_id, uint32_t device_id, clb_controls *out);
clb_status clb_controls_destroy(clb_controls ctl);
/* show_devices (src/common/env.cpp:38-56) as data */
clb_status clb_device_info_get(uint32_t device_id, clb_device_info *info);
clb_status clb_controls_sync(clb_controls ctl);
clb_status clb_controls_max_wg_size(clb_controls ctl, uint32_t *out);
clb_status clb_get_tuning(clb_controls ctl, clb_tuning *out);
clb_status clb_set_tuning(clb_controls ctl, const clb_tuning *t);
clb_status clb_last_stats(clb_controls ctl, clb_spgemm_stats *out);
/* Drop the device allocation cache of this controls' device. */
clb_status clb_pool_trim(clb_controls ctl);
/* Diagnostics of the device's buffer pool: out[0] bytes cached, out[1]
 * hipMalloc calls so far, out[2] bytes they allocated, out[3] times the cache
 * was dropped after a failed hipMalloc, out[4] free device memory now (as
 * hipMemGetInfo reports it; cached blocks count as used), out[5] device
 * memory in total. */
clb_status clb_pool_stats(clb_controls ctl, uint64_t out[6]);

/* ---- buffers (replace cl::Buffer) --------------------------------------- */
/* A zero-size buffer fails with CLBOOL_CREATE_BUFFER_ERROR, as cl::Buffer
 * does (test/test_errors.cpp:38-42). */
clb_status clb_buffer_create(clb_controls ctl, uint64_t bytes, clb_buffer *out);
clb_status clb_buffer_from_host(clb_controls ctl, const void *src, uint64_t bytes, clb_buffer *out);
clb_status clb_buffer_read(clb_controls ctl, clb_buffer buf, uint64_t offset, uint64_t bytes, void *dst);
clb_status clb_buffer_write(clb_controls ctl, clb_buffer buf, uint64_t offset, uint64_t bytes, const void *src);
clb_status clb_buffer_copy(clb_controls ctl, clb_buffer src, uint64_t src_offset, clb_buffer dst,
                           uint64_t dst_offset, uint64_t bytes);
clb_status clb_buffer_retain(clb_buffer buf);
clb_status clb_buffer_release(clb_buffer buf);
void *clb_buffer_device_ptr(clb_buffer buf);
uint64_t clb_buffer_size(clb_buffer buf);

/* ---- matrices ------------------------------------------------------------ */
clb_status clb_dcsr_release(clb_dcsr *m);
clb_status clb_coo_release(clb_coo *m);
clb_status clb_csr_release(clb_csr *m);

/* matrix_coo(controls, rows*, cols*, nrows, ncols, nnz, sorted, noDuplicates)
 * (src/core/matrix_coo.cpp:17-47): upload, sort if !sorted, drop adjacent
 * duplicates if !no_duplicates. */
clb_status clb_coo_from_host(clb_controls ctl, const uint32_t *rows, const uint32_t *cols, uint32_t nrows,
                             uint32_t ncols, uint32_t nnz, int sorted, int no_duplicates, clb_coo *out);
/* matrix_coo(controls, nrows, ncols, nnz, rows&, cols&, sorted, noDuplicates)
 * (src/core/matrix_coo.cpp:50-68). */
clb_status clb_coo_from_buffers(clb_controls ctl, uint32_t nrows, uint32_t ncols, uint32_t nnz, clb_buffer rows,
                                clb_buffer cols, int sorted, int no_duplicates, clb_coo *out);
/* coo::sort_arrays (src/coo/coo_initialization.cpp:6-32): in-place
 * lexicographic sort of (row, col) pairs. */
clb_status clb_coo_sort_arrays(clb_controls ctl, clb_buffer rows, clb_buffer cols, uint32_t n);
/* matrix_dcsr_from_cpu (src/common/matrices_conversions.cpp:93-108), with
 * explicit nrows/ncols (the reference sets both to `size`). */
clb_status clb_dcsr_from_host(clb_controls ctl, const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols,
                              uint32_t nzr, uint32_t nnz, uint32_t nrows, uint32_t ncols, clb_dcsr *out);
/* csr_from_cpu (src/common/matrices_conversions.cpp:199-208) */
clb_status clb_csr_from_host(clb_controls ctl, const uint32_t *rpt, const uint32_t *cols, uint32_t nrows,
                             uint32_t ncols, uint32_t nnz, clb_csr *out);

/* ---- conversions (src/common/matrices_conversions.hpp:9-26) -------------- */
clb_status clb_coo_to_dcsr_shallow(clb_controls ctl, const clb_coo *a, clb_dcsr *out);  /* :78-91  */
clb_status clb_dcsr_to_coo_shallow(clb_controls ctl, const clb_dcsr *a, clb_coo *out);  /* :43-57  */
clb_status clb_dcsr_to_coo_deep(clb_controls ctl, const clb_dcsr *a, clb_coo *out);     /* :59-76  */
clb_status clb_dcsr_to_csr(clb_controls ctl, const clb_dcsr *a, clb_csr *out);          /* :247-271 */
clb_status clb_csr_to_dcsr(clb_controls ctl, const clb_csr *a, clb_dcsr *out);          /* :211-244 */

/* ---- operations ----------------------------------------------------------- */
/* dcsr::matrix_multiplication_hash (src/dcsr/dcsr_matrix_multiplication_hash.cpp:51-119):
 * C = A * B over (OR, AND); throws CLBOOL_INVALID_ARGUMENT on a dimension
 * mismatch; empty inputs give an empty (a.nrows x b.ncols) result. */
clb_status clb_dcsr_spgemm_hash(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out);
/* The same product with a mode (plain two-pass / learned one-pass), an
 * optional row block of A, and statistics.  Output is bit-identical across
 * modes. */
clb_status clb_dcsr_spgemm(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, const clb_spgemm_opts *opts,
                           clb_dcsr *out, clb_spgemm_stats *stats);
/* dcsr::matrix_multiplication (src/dcsr/dcsr_matrix_multiplication.cpp:26-168):
 * the same product by expand-sort-compress (rows sorted in registers, in
 * LDS or, above 2^13 products, as 64-bit keys in HBM), with no hash or
 * bitmap accumulator: an independent cross-check of the hash path.  Same
 * errors and empty-result shapes as clb_dcsr_spgemm_hash. */
clb_status clb_dcsr_spgemm_esc(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out);

/* ---- a product taken in row blocks (no counterpart in the reference: its
 *      matrix_multiplication_hash wraps around past 2^32 - 1 nonzeros,
 *      hash.cpp:312,318).  The plan runs each mode's sizing pass over all of
 *      A's stored rows and cuts them into consecutive blocks whose results
 *      hold at most block_bound nonzeros (0 or >= 2^32 - 1: 2^32 - 1):
 *      PLAIN by exact sizes from the symbolic pass (a block then runs the
 *      numeric pass only), LEARNED by predicted capacities, BOUND by
 *      min(products, ncols).  A LEARNED block whose exact result exceeds the
 *      bound is cut again by the exact sizes it found and run again, so
 *      nblocks can grow during the runs: read it again after each run.  The
 *      plan keeps references to A and B.  Block k is stored rows
 *      [row_begin, row_end) of A, and its result is the same as
 *      clb_dcsr_spgemm over that row block. */
typedef struct clb_spgemm_plan_s *clb_spgemm_plan;
typedef struct clb_spgemm_plan_info {
    uint32_t mode;
    uint32_t nblocks;      /* current number of blocks                         */
    uint32_t replans;      /* blocks cut again after exceeding the bound       */
    uint32_t reserved;
    uint64_t products;     /* intermediate products of the whole product       */
    uint64_t planned_nnz;  /* sum of the sizes the cut used                    */
    double   ms_plan;      /* host time of clb_spgemm_plan_create              */
} clb_spgemm_plan_info;
clb_status clb_spgemm_plan_create(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, uint32_t mode,
                                  uint64_t block_bound, clb_spgemm_plan *out);
clb_status clb_spgemm_plan_info_get(clb_spgemm_plan plan, clb_spgemm_plan_info *info);
clb_status clb_spgemm_plan_block(clb_spgemm_plan plan, uint32_t k, uint32_t *row_begin, uint32_t *row_end);
clb_status clb_spgemm_plan_run(clb_controls ctl, clb_spgemm_plan plan, uint32_t k, uint32_t timing, clb_dcsr *out,
                               clb_spgemm_stats *stats);
/* Runs every block in order and hands each one to fn as soon as it is
 * complete: fn(user, k, &block, &stats) owns the block (it releases it with
 * clb_dcsr_release) and returns 0 to go on, nonzero to stop.  In the one-pass
 * modes block k + 1's passes start while block k's last phase (compaction and
 * dense-row emission) still runs, and block k + 1's row kernels are launched
 * before block k's sizes are read back, so the blocks overlap (each when
 * device memory allows); fn is called on the calling thread.  Results are the
 * same as clb_spgemm_plan_run's. */
typedef int32_t (*clb_spgemm_block_fn)(void *user, uint32_t k, clb_dcsr *block, const clb_spgemm_stats *stats);
clb_status clb_spgemm_plan_run_all(clb_controls ctl, clb_spgemm_plan plan, uint32_t timing, clb_spgemm_block_fn fn,
                                   void *user);
clb_status clb_spgemm_plan_release(clb_spgemm_plan plan);

/* dcsr::kronecker_product (src/dcsr/dcsr_kronecker_product.cpp:6-65) */
clb_status clb_dcsr_kronecker(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out);
/* dcsr::transpose (src/dcsr/dcsr_transpose.cpp:5-17): an empty input gives
 * an empty (a.ncols x a.nrows) result. */
clb_status clb_dcsr_transpose(clb_controls ctl, const clb_dcsr *a, clb_dcsr *out);
/* dcsr::submatrix (src/dcsr/dcsr_submatrix.cpp:98-156): rows [i, i+nrows),
 * columns [j, j+ncols), shifted to the origin; shape nrows x ncols.  As in the
 * reference, an empty input or an empty result gives a 0 x 0 matrix. */
clb_status clb_dcsr_submatrix(clb_controls ctl, const clb_dcsr *a, uint32_t i, uint32_t j, uint32_t nrows,
                              uint32_t ncols, clb_dcsr *out);
/* dcsr::reduce (src/dcsr/dcsr_reduce.cpp:5-31): column 0 set in every
 * nonempty row; shape a.nrows x 1 (the reference labels it a.nzr x 1).  An
 * empty input gives a 0 x 0 matrix. */
clb_status clb_dcsr_reduce(clb_controls ctl, const clb_dcsr *a, clb_dcsr *out);
/* A 64-bit digest of a DCSR matrix computed on the device: the sum over its
 * nonzeros of a mix of (position, row id, column), combined with its shape
 * and sizes.  Equal matrices have equal digests; a changed, moved, missing or
 * extra entry changes it except with probability about 2^-64.  No
 * counterpart in the reference: it replaces reading matrices back to compare
 * them on the host (compare_matrices, src/common/utils.cpp:42-71). */
clb_status clb_dcsr_digest(clb_controls ctl, const clb_dcsr *a, uint64_t *digest);
/* coo::matrix_addition (src/coo/coo_matrix_addition.cpp:32-73) */
clb_status clb_coo_addition(clb_controls ctl, const clb_coo *a, const clb_coo *b, clb_coo *out);
/* csr::matrix_addition (src/csr/csr_addition.cpp:26-168) */
clb_status clb_csr_addition(clb_controls ctl, const clb_csr *a, const clb_csr *b, clb_csr *out);
/* DCSR union C = A | B (the X += T step of a closure; row-wise merge). */
clb_status clb_dcsr_addition(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, clb_dcsr *out);

/* ---- closure step of CFPQ/RPQ evaluation (no counterpart in the reference;
 *      its caller, spbla, loops over matrix_multiplication_hash and additions) */
#define CLB_CLOSURE_PER_TRIPLE 0u  /* one SpGEMM and one union per triple        */
#define CLB_CLOSURE_BATCHED    1u  /* all products of the step as SpGEMMs on
                                      block-stacked operands (one per row block
                                      of the stack), one union per X           */
typedef struct clb_closure_opts {
    uint32_t mode;         /* CLB_CLOSURE_*                                   */
    uint32_t spgemm_mode;  /* CLB_SPGEMM_PLAIN or CLB_SPGEMM_LEARNED           */
    uint32_t block_bound;  /* batched: cut the stacked product into row blocks
                              whose output bound (sum over the block's rows of
                              min(products, n)) is at most this (0: 2^32 - 1;
                              values below 2n are raised to 2n)           */
    uint32_t reserved;
} clb_closure_opts;
typedef struct clb_closure_stats {
    uint64_t products;     /* intermediate products of the step               */
    uint64_t added;        /* nonzeros added to the X matrices                */
    uint32_t changed;      /* matrices that grew (0: fixed point reached)     */
    uint32_t spgemm_calls;
    double   ms;
    uint64_t max_block_bound; /* batched: largest output bound of a row block  */
    double   ms_products;  /* host time of the SpGEMMs (stacking included)   */
    double   ms_unions;    /* host time of the unions                        */
    uint64_t fallback_rows; /* learned mode: rows re-run exactly, all blocks */
    uint64_t dense_rows;   /* rows emitted from dumped bitmaps, all blocks   */
} clb_closure_stats;
/* One step X_t |= Y_t * Z_t for every triple t = (X, Y, Z) (indices into
 * mats, triples[3t..3t+2]); all products use the matrices as they were at
 * the start of the step.  mats[] is updated in place: every matrix that grows
 * is released and replaced by the new one.  All matrices must be n x n. */
clb_status clb_closure_step(clb_controls ctl, clb_dcsr *mats, uint32_t nmats, const uint32_t *triples,
                            uint32_t ntriples, const clb_closure_opts *opts, clb_closure_stats *stats);

/* ---- learned row-size predictor (no counterpart in the reference) -------- */
/* An MLP 5 -> 32 -> 32 -> 1 (fp32, ReLU) maps per-row features (log2 of
 * 1 + products, A-row length, min and max B-row length, column span, each
 * divided by 32) to log2(1 + nnz of the output row) / 32.  Parameters are
 * CLB_MLP_PARAMS floats laid out W1[32][5], b1[32], W2[32][32], b2[32],
 * w3[32], b3.  Training data are exact row sizes harvested from the plain
 * symbolic pass; training runs Adam on the GPU with the library's own
 * forward/backward kernels. */
#define CLB_MLP_PARAMS 1281
typedef struct clb_train_opts {
    uint32_t epochs;
    uint32_t batch;
    float    lr;
    uint32_t seed;
} clb_train_opts;
typedef struct clb_train_stats {
    double   loss_first;  /* mean 0.5*(y-t)^2 of the first epoch */
    double   loss_last;   /* ... of the last epoch               */
    uint64_t samples;
    uint32_t steps;
    double   ms;
} clb_train_stats;
/* Initialise the parameters (seeded He initialisation) and clear the data. */
clb_status clb_predictor_reset(clb_controls ctl, uint32_t seed);
/* Append (features, exact row size) of up to max_rows rows of A * B. */
clb_status clb_predictor_harvest(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, uint32_t max_rows,
                                 uint64_t *total_samples);
clb_status clb_predictor_train(clb_controls ctl, const clb_train_opts *opts, clb_train_stats *stats);
clb_status clb_predictor_get_params(clb_controls ctl, float *params, uint32_t n);
clb_status clb_predictor_set_params(clb_controls ctl, const float *params, uint32_t n);
/* Predicted output size of every stored row of A * B (a.nzr floats). */
clb_status clb_predictor_predict(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, float *pred_nnz);
/* The MLP kernels on host data, for checking them against a CPU oracle:
 * forward (y[n]), mean loss and its gradient (grad[CLB_MLP_PARAMS]), and one
 * Adam step (in place on params, m, v). */
clb_status clb_mlp_forward(clb_controls ctl, const float *params, const float *x, uint32_t n, float *y);
clb_status clb_mlp_loss_grad(clb_controls ctl, const float *params, const float *x, const float *t, uint32_t n,
                             float *grad, float *loss);
clb_status clb_mlp_adam(clb_controls ctl, float *params, float *m, float *v, const float *grad, float lr, float b1,
                        float b2, float eps, uint32_t step);

/* ---- primitives (src/common/cl_operations.hpp:9-17) ---------------------- */
/* prefix_sum (src/common/cl_operations.cpp:4-88): in-place exclusive scan of
 * n uint32 values; *total receives the sum. */
clb_status clb_prefix_sum(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t *total);
/* clb_prefix_sum without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * tile sums live in a fixed 256 KiB device array per GPU, so n is at most
 * 2^28 (larger n: CLBOOL_INVALID_ARGUMENT).  Calls that overlap on one GPU
 * share that array and race. */
clb_status clb_prefix_sum_noalloc(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t *total);
/* fill_with (src/common/cl_operations.cpp:90-100) */
clb_status clb_fill_with(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t value);

/* KernelCache::get_kernel (src/core/kernel_cache.hpp:46-56): kernels are
 * compiled ahead of time, so this only checks that `program`/`kernel` names
 * one of them; unknown names fail with CLBOOL_NO_SUCH_PROGRAM or
 * CLBOOL_CREATE_KERNEL_ERROR as in the reference. */
clb_status clb_kernel_lookup(const char *program, const char *kernel);

#ifdef __cplusplus
}
#endif
#endif /* CLBOOL_C_H */
