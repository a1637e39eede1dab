// rb_mfma.h -- MFMA operand types and the lane layout of the 32x32 output tile (gfx950).
//
// v_mfma_f32_32x32x16_{f16,bf16}: D(32x32) += A(32x16) * B(16x32).  Lane l holds
//   A[m = l%32][k-slot (l/32, j)], j = 0..7      (8 halves)
//   B[k-slot (l/32, j)][n = l%32], j = 0..7      (8 halves)
//   D[row(r, l)][n = l%32], r = 0..15            (16 floats)
// The k-slots of A and B are paired by the hardware, so any one-to-one map of (l/32, j) to
// k serves both operands; here k = 8*(l/32) + j.  The output row of register r is
// row = 8*(r/4) + 4*(l/32) + r%4 (checked on the device by rb_probe_mfma_layout).
#pragma once
#include <hip/hip_runtime.h>

namespace rb {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef double doublex4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int mfma32_row(int r, int lane) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

}  // namespace rb
