// sort.hpp -- pair-key sort and union helpers shared by the COO/CSR/DCSR
// operations.
#pragma once

#include "clb_internal.hpp"

namespace clb {

// Stable LSD radix sort of 64-bit keys row << 32 | col, over the low
// col_bits of the column half and the low row_bits of the row half.
void radix_sort_keys(clb_controls_s *ctl, u64 *keys, u64 n, int col_bits, int row_bits);
void sort_pair_keys(clb_controls_s *ctl, u64 *keys, u64 n, u32 max_row, u32 max_col);
void pack_pairs(clb_controls_s *ctl, const u32 *rows, const u32 *cols, u64 n, u64 *keys);
void unpack_pairs(clb_controls_s *ctl, const u64 *keys, u64 n, u32 *rows, u32 *cols);
void key_bounds(clb_controls_s *ctl, const u64 *keys, u64 n, u32 *max_row, u32 *max_col);
// Sorted keys -> new deduplicated (rows, cols) buffers; returns their length.
u32 unique_keys_to_pairs(clb_controls_s *ctl, const u64 *keys, u64 n, clb_buffer_s **rows_out,
                         clb_buffer_s **cols_out);
// Merge of two ascending key lists into out[0, na + nb) (merge path, kept
// duplicates).
void merge_keys(clb_controls_s *ctl, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out);
// Sorted keys (row << 32 | col, duplicates allowed) -> a new canonical DCSR
// of their distinct pairs; out.nrows/ncols are kept as given.
void keys_to_dcsr(clb_controls_s *ctl, const u64 *keys, u64 n, clb_dcsr &out);
void coo_finish(clb_controls_s *ctl, clb_buffer_s *rows, clb_buffer_s *cols, u32 nrows, u32 ncols, u32 nnz,
                bool sorted, bool nodup, clb_coo &out);

// COO rows -> DCSR rpt/rows (create_rows_pointers, matrices_conversions.cpp:7-37).
void rows_to_dcsr(clb_controls_s *ctl, const u32 *rows, u32 nnz, clb_buffer_s **rpt_out, clb_buffer_s **rows_out,
                  u32 *nzr_out);
// DCSR rpt/rows -> per-nonzero row ids (dscr_to_coo.cl:8-23).
void dcsr_expand_rows(clb_controls_s *ctl, const u32 *rpt, const u32 *rows, u32 nzr, u32 nnz, u32 *out_rows);

}  // namespace clb
