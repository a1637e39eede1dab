// rb_fused_mc8.hip -- the plain Monte Carlo fused kernel with eight groups per pass (for units
// with 7, 8 or 13-16 maturity groups, where it saves a pass over the accumulators); the stages
// are described in rb_fused_mc.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc8(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 8>(mode, upw, a, grid, st);
}

}  // namespace rb
