// rb_fused_rt5.hip -- the Romano-Touzi fused kernel with five groups per pass, for units of five
// maturity groups (no dead group slot in the row loop) or nine to ten (two passes); the stages are
// described in rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_rt5(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, 5>(mode, upw, a, grid, st);
}

}  // namespace rb
