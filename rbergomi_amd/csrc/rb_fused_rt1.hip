// rb_fused_rt1.hip -- the Romano-Touzi fused kernel with one group per pass, for units of a single
// maturity group (single-maturity calibration requests); the stages are described in
// rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_rt1(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, 1>(mode, upw, a, grid, st);
}

}  // namespace rb
