// rb_fused_rt3.hip -- the Romano-Touzi fused kernel with three groups per pass, for units of three
// maturity groups; the stages are described in rb_fused.hip.
#include "rb_fused_kernel.h"

namespace rb {

hipError_t launch_fused_rt3(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, 3>(mode, upw, a, grid, st);
}

}  // namespace rb
