// rb_fused_mc1.hip -- the plain Monte Carlo fused kernel with one group per pass, for units of a single
// maturity group (single-maturity calibration requests); the stages are described in
// rb_fused_mc.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc1(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 1>(mode, upw, a, grid, st);
}

}  // namespace rb
