// rb_fused_mc2.hip -- the plain Monte Carlo fused kernel with two groups per pass, for units of two
// maturity groups; the stages are described in rb_fused_mc.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc2(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 2>(mode, upw, a, grid, st);
}

}  // namespace rb
