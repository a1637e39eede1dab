// rb_fused_rt2.hip -- the Romano-Touzi fused kernel with two groups per pass, for units of two
// maturity groups (no dead group slot in the row loop); the stages are described in rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_rt2(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, 2>(mode, upw, a, grid, st);
}

}  // namespace rb
