// rb_fused_mc.hip -- the fused kernel for the plain Monte Carlo estimator (computeIV_tensor), six
// groups per pass.
//
// Same stages as the Romano-Touzi kernel (rb_fused.hip) up to the integrals; a third Gaussian
// stream Wperp (stream 2 of the counter space, as on the FP64 path) drives the price:
// ln S_T = sum_j sqrt(v_j) sqrt(dt) (rho W1_j + sqrt(1 - rho^2) Wperp_j) - dt v_j / 2
// (updateZ/scaleZ/updateS, rBergLib/rBergomiMT.cpp:111-136), summed per (path, group) as
// rho Is + sqrt(1 - rho^2) Ip - Iv / 2 with Ip = sqrt(dt) sum sqrt(v_j) Wperp_j, and the payoff is
// (S_T - K)^+ (posPart, :264-266) in FP64.
#include "rb_fused_kernel.h"

namespace rb {

// six groups per pass (the one-, two-, four-, five- and eight-group variants are in their own
// files, built in parallel)
hipError_t launch_fused_mc6(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<true, 6>(mode, upw, a, grid, st);
}

}  // namespace rb
