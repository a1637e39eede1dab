// rb_probe.hip -- device probe of the MFMA operand/output layout.
#include <hip/hip_runtime.h>

#include "rb_fused.h"
#include "rb_kernels.h"
#include "rb_mfma.h"

namespace rb {

// One wave computes D = A(32x16) * B(16x32) with the operand layout of rb_mfma.h
// (A slot (h, j) = A[m][8h+j], B slot (h, j) = B[8h+j][n]) and writes the raw accumulator
// registers: out[lane*16 + r].  Small integer inputs make every product and sum exact.
// mode 0: one v_mfma_f32_32x32x16_f16; 1: the bf16 form; 2: eight v_mfma_f32_32x32x2_f32, the
// j-th on slot j of each half (as the FP32-operand fused kernel pairs them).
// mode 3: one v_mfma_f64_16x16x4_f64 of A[16][4] (A[m*4 + k]) by B[4][16] (B[k*16 + n]) with lane
// l supplying A[l%16][l/16] and B[l/16][l%16]; out[lane*4 + r] = accumulator register r.
__global__ void k_probe_mfma_layout(const float* __restrict__ A, const float* __restrict__ B,
                                    float* __restrict__ out, int mode) {
    const int l = threadIdx.x;
    const int m = l & 31, h = l >> 5;
    if (mode == 3) {
        doublex4_t d = {0.0, 0.0, 0.0, 0.0};
        d = __builtin_amdgcn_mfma_f64_16x16x4f64(double(A[(l & 15) * 4 + (l >> 4)]),
                                                 double(B[(l >> 4) * 16 + (l & 15)]), d, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[l * 4 + r] = float(d[r]);
        return;
    }
    floatx16_t acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    if (mode == 2) {
        for (int j = 0; j < 8; ++j)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(A[m * 16 + 8 * h + j], B[(8 * h + j) * 32 + m], acc, 0, 0, 0);
    } else if (mode == 1) {
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

hipError_t launch_probe_mfma_layout(const float* A, const float* B, float* out, int mode,
                                    hipStream_t st) {
    hipLaunchKernelGGL(k_probe_mfma_layout, dim3(1), dim3(64), 0, st, A, B, out, mode);
    return hipGetLastError();
}

// Mixed-precision payoff of the fused kernel for given (ln S', sigma, K): out[i]
__global__ void k_probe_payoff(const double* __restrict__ lnS, const double* __restrict__ sig,
                               const double* __restrict__ K, int n, double* __restrict__ out,
                               float* __restrict__ mills_out, const float* __restrict__ z, int form) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PathPG q;
    q.lnS = lnS[i];
    q.sig = sig[i];
    pathpg_f32(q);
    const double lnK = log(K[i]);
    const OptF of = opt_f32(lnK, K[i]);
    out[i] = form ? payoff_mixed<true>(q, of, lnK, K[i]) : payoff_mixed<false>(q, of, lnK, K[i]);
    mills_out[i] = mills_f32(z[i]);
}

hipError_t launch_probe_payoff(const double* lnS, const double* sig, const double* K, int n,
                               double* out, float* mills_out, const float* z, int form, hipStream_t st) {
    hipLaunchKernelGGL(k_probe_payoff, dim3((n + 255) / 256), dim3(256), 0, st, lnS, sig, K, n, out,
                       mills_out, z, form);
    return hipGetLastError();
}

// Plain streaming copy (16-byte vectors): the known-size reference kernel for the HBM
// traffic counters, and a measure of the achievable copy bandwidth.
__global__ void k_probe_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, long n16) {
    for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n16; i += long(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_probe_copy(const void* src, void* dst, long bytes, hipStream_t st) {
    const long n16 = bytes / 16;
    hipLaunchKernelGGL(k_probe_copy, dim3(4096), dim3(256), 0, st, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    return hipGetLastError();
}

// Issue-rate probe: eight independent chains per lane of one instruction class, 2048
// workgroups of 256 threads (eight waves per SIMD), so the SIMDs issue back to back.  op 0: FP32
// FMA, 1: FP32 exp2 (transcendental), 2: 32 x 32 -> 64-bit integer multiply-add plus a 32-bit add
// (the Philox round's product), 3: FP64 FMA.
constexpr int PROBE_ISSUE_IT = 1024;

template <int OP>
__global__ void __launch_bounds__(256) k_probe_issue(float* __restrict__ out, uint32_t seed) {
    float x[8];
    double y[8];
    uint32_t u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        x[j] = 0.3f + 0.01f * float(j) + 1e-6f * float(threadIdx.x);
        y[j] = double(x[j]);
        u[j] = seed + 77u * uint32_t(j) + threadIdx.x;
    }
    for (int i = 0; i < PROBE_ISSUE_IT; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if constexpr (OP == 0) x[j] = fmaf(x[j], 0.999f, 0.001f);
            if constexpr (OP == 1) x[j] = __builtin_amdgcn_exp2f(-x[j]);
            if constexpr (OP == 2) {
                const uint64_t q = uint64_t(u[j]) * 0xD2511F53u;
                u[j] = uint32_t(q >> 32) + uint32_t(q);
            }
            if constexpr (OP == 3) y[j] = fma(y[j], 0.999, 0.001);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += x[j] + float(y[j]) + float(u[j] & 1023u);
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

hipError_t launch_probe_issue(int op, float* out, uint32_t seed, hipStream_t st) {
    const dim3 g(2048), b(256);
    switch (op) {
        case 0: hipLaunchKernelGGL(k_probe_issue<0>, g, b, 0, st, out, seed); break;
        case 1: hipLaunchKernelGGL(k_probe_issue<1>, g, b, 0, st, out, seed); break;
        case 2: hipLaunchKernelGGL(k_probe_issue<2>, g, b, 0, st, out, seed); break;
        case 3: hipLaunchKernelGGL(k_probe_issue<3>, g, b, 0, st, out, seed); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// wave64 instructions of one class per probe launch and SIMD (loop body only)
double probe_issue_insts_per_simd(int simds) { return 2048.0 * 4 * PROBE_ISSUE_IT * 8 / simds; }

}  // namespace rb
