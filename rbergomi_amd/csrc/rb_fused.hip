// rb_fused.hip -- fused mixed-precision pricing kernel for N <= 1024 (gfx950).
//
// One workgroup (4 waves) owns a tile of TP = 32*PB paths at a time and runs the whole path
// of the reference for it without touching HBM in between:
//   1. Philox4x32-10 Gaussians W1 (FP32 Box-Muller) -> LDS in the B-operand fragment layout of
//      v_mfma_f32_32x32x{16_f16,16_bf16,2_f32}: FP16 mode rounds W1 to FP16 (the value every
//      later stage uses); BF16 mode splits it into hi + lo; FP32 mode keeps it (replaces
//      genGaussianMT, rBergLib/rBergomiMT.cpp:169-172)
//   2. Y' = L W1 on the matrix cores, L the lower-triangular Toeplitz matrix of Gamma with the
//      first-cell weight on its diagonal, L split into hi + lo: L_hi W1 + L_lo W1 (FP16: two
//      products, exact up to the FP32 accumulation), L_hi W1_hi + L_hi W1_lo + L_lo W1_hi
//      (BF16: three), or one FP32 product (replaces the FFTW convolution of updateWtilde,
//      :83-109).
//      The A operand of a
//      32x16 block depends only on its diagonal offset d, so A fragments come from a small
//      per-H table (no N x N matrix).  A wave walks d in phases over which the set of its
//      active row blocks is fixed, so the MFMAs of different row blocks interleave.
//   3. Epilogue in registers: Wtilde = sc (Y' + b W1perp) with W1perp regenerated from the
//      same counters, then for every (H, eta, T) group sqrt(v_{i+1}) = exp2(c1 Wtilde_i -
//      al P_{i+1} - be) and the sums of v and sqrt(v) W1 (scaleWtilde/updateV/intVdt/
//      intRootVdW, :73-80, 163-167, 427-436): FP32 terms summed per 16-row block, FP64 beyond.
//      The row terms al P_{i+1} + be come from a per-group table (path-independent).
//   4. Per (path, group, rho): ln S', sigma, 1/sigma, S' once; per strike the conditioned
//      Black-Scholes payoff S' Phi(d1) - K Phi(d2) (:383-391) in FP32, with the FP64 normal
//      CDF for deep out-of-the-money lanes (d1 < -8), for the options of up to 8 groups at a
//      time.  The paths of an option are split over S = 2^k lanes (S ~ 256 / options) and summed
//      in FP64 with a fixed butterfly into the register of the option's owner lane.
#include <hip/hip_runtime.h>

#include "rb_fused_kernel.h"

namespace rb {

size_t fused_req_lds_bytes(const FusedReq& a) {
    return size_t(a.PB) * a.NK * a.chunk_u4 * 16 + size_t(4) * a.gchunk * 32 * (a.mc ? 3 : 2) * 8 +
           size_t(a.TP) * a.pgmax * sizeof(PathPG) + (a.acc_in_lds ? size_t(a.out_size) * 2 * 8 : 0);
}

size_t fused_lds_bytes(const FusedBatchArgs& b) {
    return size_t(b.bf_u4) * 16 + size_t(b.wpart_d) * 8 + size_t(b.pgd_n) * sizeof(PathPG) +
           size_t(b.acc_d) * 8;
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_int(int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_fused_rt(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return launch_mu<false, FUSED_GREG>(mode, upw, a, grid, st);
}

hipError_t launch_fused(int mode, int upw, bool mc, int greg, const FusedBatchArgs& a, int grid,
                        hipStream_t st) {
    if (!mc) {
        if (greg == 1) return launch_fused_rt1(mode, upw, a, grid, st);
        if (greg == 2) return launch_fused_rt2(mode, upw, a, grid, st);
        if (greg == 3) return launch_fused_rt3(mode, upw, a, grid, st);
        if (greg == 4) return launch_fused_rt4(mode, upw, a, grid, st);
        if (greg == 5) return launch_fused_rt5(mode, upw, a, grid, st);
        return greg == 8 ? launch_fused_rt(mode, upw, a, grid, st) : launch_fused_rt6(mode, upw, a, grid, st);
    }
    if (greg == 1) return launch_fused_mc1(mode, upw, a, grid, st);
    if (greg == 2) return launch_fused_mc2(mode, upw, a, grid, st);
    if (greg == 3) return launch_fused_mc3(mode, upw, a, grid, st);
    if (greg == 4) return launch_fused_mc4(mode, upw, a, grid, st);
    if (greg == 5) return launch_fused_mc5(mode, upw, a, grid, st);
    return greg == 8 ? launch_fused_mc8(mode, upw, a, grid, st) : launch_fused_mc6(mode, upw, a, grid, st);
}

}  // namespace rb
