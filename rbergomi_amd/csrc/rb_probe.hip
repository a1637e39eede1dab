// rb_probe.hip -- device probe of the MFMA operand/output layout.
#include <hip/hip_runtime.h>

#include "rb_kernels.h"
#include "rb_mfma.h"

namespace rb {

// One wave computes D = A(32x16) * B(16x32) with the operand layout of rb_mfma.h
// (A slot (h, j) = A[m][8h+j], B slot (h, j) = B[8h+j][n]) and writes the raw accumulator
// registers: out[lane*16 + r].  Small integer inputs make every product and sum exact.
__global__ void k_probe_mfma_layout(const float* __restrict__ A, const float* __restrict__ B,
                                    float* __restrict__ out, int bf16) {
    const int l = threadIdx.x;
    const int m = l & 31, h = l >> 5;
    floatx16_t acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    if (bf16) {
        bf16x8_t a, b;
        for (int j = 0; j < 8; ++j) {
            a[j] = (__bf16)A[m * 16 + 8 * h + j];
            b[j] = (__bf16)B[(8 * h + j) * 32 + m];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    } else {
        half8_t a, b;
        for (int j = 0; j < 8; ++j) {
            a[j] = (_Float16)A[m * 16 + 8 * h + j];
            b[j] = (_Float16)B[(8 * h + j) * 32 + m];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
    }
    for (int r = 0; r < 16; ++r) out[l * 16 + r] = acc[r];
}

hipError_t launch_probe_mfma_layout(const float* A, const float* B, float* out, int bf16,
                                    hipStream_t st) {
    hipLaunchKernelGGL(k_probe_mfma_layout, dim3(1), dim3(64), 0, st, A, B, out, bf16);
    return hipGetLastError();
}

}  // namespace rb
