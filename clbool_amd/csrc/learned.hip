// learned.hip -- learned row-size predictor and the one-pass SpGEMM (stub).
#include "spgemm.hpp"

namespace clb {
namespace spg {
void spgemm_learned(clb_controls_s *, const clb_dcsr &, const clb_dcsr &, u32, u32, bool, clb_dcsr &,
                    clb_spgemm_stats &) {
    raise(INVALID_VALUE, "learned mode is not available in this build");
}
}  // namespace spg
}  // namespace clb

namespace clb {
void learned_destroy_state(clb_controls_s *) {}
}  // namespace clb
