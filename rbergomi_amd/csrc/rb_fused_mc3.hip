// rb_fused_mc3.hip -- the plain Monte Carlo fused kernel with three groups per pass, for units of three
// maturity groups; the stages are described in rb_fused_mc.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc3(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 3>(mode, upw, a, grid, st);
}

}  // namespace rb
