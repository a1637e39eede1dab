// rb_fused_mc5.hip -- the plain Monte Carlo fused kernel with five groups per pass, for units of
// five maturity groups or nine to ten (two passes); the stages are described in rb_fused_mc.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc5(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 5>(mode, upw, a, grid, st);
}

}  // namespace rb
