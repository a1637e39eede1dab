// rb_fused_rt6.hip -- the Romano-Touzi fused kernel with six groups per pass (units of six
// maturity groups, or eleven to twelve in two passes: fewer dead group slots per row and fewer
// accumulator registers than the eight-group variant in rb_fused.hip); the stages are described
// in rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_rt6(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, 6>(mode, upw, a, grid, st);
}

}  // namespace rb
