// rb_fused_mc4.hip -- the plain Monte Carlo fused kernel with four groups per pass, for units of four
// maturity groups; the stages are described in rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_mc4(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 4>(mode, upw, a, grid, st);
}

}  // namespace rb
