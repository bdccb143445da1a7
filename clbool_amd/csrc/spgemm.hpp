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
    DevArray<u32> prod;
    DevArray<u64> scal;
    Mats g{};
    BinCfg cfg{};
};

BinCfg make_bincfg(const clb_controls_s *ctl, u32 ncols);
void prepare(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 nrows, Setup &s);
void bin_rows(clb_controls_s *ctl, const Setup &s, const u32 *key, u32 *nnz_row_zero, DevArray<u32> &perm,
              DevArray<u32> &binbuf, u32 hist_host[NBINS + 1], u64 binprod_host[NBINS], bool readback_now);
void launch_bins(clb_controls_s *ctl, const Setup &s, const u32 *perm, const u32 off[NBINS + 1], bool numeric,
                 u32 *nnz_row, const u64 *out_off, u32 *out_cols, u32 *overflow);
void scan_sizes(clb_controls_s *ctl, Setup &s, const u32 *nnz_row, DevArray<u64> &off, DevArray<u32> &pos);
void assemble(clb_controls_s *ctl, const Setup &s, const clb_dcsr &a, const u32 *nnz_row, const u64 *off,
              const u32 *pos, u64 nnz, u32 nzr, clb_buffer_s *cols, clb_dcsr &out);
void check_dims(const clb_dcsr &a, const clb_dcsr &b);

void spgemm_plain(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                  bool timing, clb_dcsr &out, clb_spgemm_stats &st);
// learned.hip
void spgemm_learned(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 row_begin, u32 row_end,
                    bool timing, clb_dcsr &out, clb_spgemm_stats &st);
void spgemm_dispatch(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, const clb_spgemm_opts &o,
                     clb_dcsr &out, clb_spgemm_stats &st);

}  // namespace spg
}  // namespace clb
