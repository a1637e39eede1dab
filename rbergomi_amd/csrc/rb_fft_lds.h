// rb_fft_lds.h -- complex FP32 transforms of a power-of-two length held in LDS, shared by the
// long-path kernels (rb_fft.hip, rb_fused_fft.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rb {

// LDS slot of element i: one pad slot after every 32 elements, so the power-of-two strides of
// the butterflies and of the bit-reversed loads spread over the banks
__device__ __forceinline__ int fpad(int i) { return i + (i >> 5); }
// LDS slots for L elements
inline size_t fft_lds_slots(int log2L) { return (size_t(1) << log2L) + (size_t(1) << log2L >> 5); }

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-LDS transforms of z (element i in slot fpad(i)) by 256 threads, two radix-2 stages per pass (radix 4: four points
// per butterfly, half the LDS passes and barriers of radix 2).  tw[k] = exp(-2 pi i k / L).
// Forward: decimation in time from bit-reversed order to natural order.  Stages s and s + 1
// (span 2^s): x1, x3 take twiddle a = W_{2 span}^pos, then x2', x3' take b = W_{4 span}^pos and
// b W_4 = -i b.
__device__ __forceinline__ void fft_dit_r4(float2* z, const float2* __restrict__ tw, int log2L, int tid) {
    const int L = 1 << log2L;
    int s = 0;
    if (log2L & 1) {
        for (int b = tid; b < (L >> 1); b += 256) {
            const float2 x = z[fpad(2 * b)], y = z[fpad(2 * b + 1)];
            z[fpad(2 * b)] = make_float2(x.x + y.x, x.y + y.y);
            z[fpad(2 * b + 1)] = make_float2(x.x - y.x, x.y - y.y);
        }
        __syncthreads();
        s = 1;
    }
    for (; s < log2L; s += 2) {
        const int span = 1 << s;
        for (int t = tid; t < (L >> 2); t += 256) {
            const int pos = t & (span - 1);
            const int i0 = ((t >> s) << (s + 2)) + pos, i1 = i0 + span, i2 = i1 + span, i3 = i2 + span;
            const float2 a = tw[pos << (log2L - s - 1)], b = tw[pos << (log2L - s - 2)];
            const float2 x0 = z[fpad(i0)], x1 = z[fpad(i1)], x2 = z[fpad(i2)], x3 = z[fpad(i3)];
            const float2 t1 = cmulf(a, x1), t3 = cmulf(a, x3);
            const float2 x0p = make_float2(x0.x + t1.x, x0.y + t1.y), x1p = make_float2(x0.x - t1.x, x0.y - t1.y);
            const float2 x2p = make_float2(x2.x + t3.x, x2.y + t3.y), x3p = make_float2(x2.x - t3.x, x2.y - t3.y);
            const float2 u = cmulf(b, x2p), v = cmulf(b, x3p);
            z[fpad(i0)] = make_float2(x0p.x + u.x, x0p.y + u.y);
            z[fpad(i2)] = make_float2(x0p.x - u.x, x0p.y - u.y);
            z[fpad(i1)] = make_float2(x1p.x + v.y, x1p.y - v.x);  // x1' + (-i) v
            z[fpad(i3)] = make_float2(x1p.x - v.y, x1p.y + v.x);
        }
        __syncthreads();
    }
}

// Inverse (unnormalised, conjugate twiddles): decimation in frequency from natural order to
// bit-reversed order.  Stages s and s - 1 (span1 = 2^(s-1)): pairs (x0, x2), (x1, x3) with
// cb = conj W_{4 span1}^pos and cb i, then (u0, u1), (u2, u3) with ca = conj W_{2 span1}^pos.
__device__ __forceinline__ void fft_dif_inv_r4(float2* z, const float2* __restrict__ tw, int log2L, int tid) {
    const int L = 1 << log2L;
    int s = log2L - 1;
    for (; s >= 1; s -= 2) {
        const int sh = s - 1, span1 = 1 << sh;
        for (int t = tid; t < (L >> 2); t += 256) {
            const int pos = t & (span1 - 1);
            const int i0 = ((t >> sh) << (s + 1)) + pos, i1 = i0 + span1, i2 = i1 + span1, i3 = i2 + span1;
            const float2 b0 = tw[pos << (log2L - s - 1)], a0 = tw[pos << (log2L - s)];
            const float2 cb = make_float2(b0.x, -b0.y), ca = make_float2(a0.x, -a0.y);
            const float2 x0 = z[fpad(i0)], x1 = z[fpad(i1)], x2 = z[fpad(i2)], x3 = z[fpad(i3)];
            const float2 u0 = make_float2(x0.x + x2.x, x0.y + x2.y);
            const float2 u2 = cmulf(make_float2(x0.x - x2.x, x0.y - x2.y), cb);
            const float2 u1 = make_float2(x1.x + x3.x, x1.y + x3.y);
            const float2 w = cmulf(make_float2(x1.x - x3.x, x1.y - x3.y), cb);
            const float2 u3 = make_float2(-w.y, w.x);  // i w
            z[fpad(i0)] = make_float2(u0.x + u1.x, u0.y + u1.y);
            z[fpad(i1)] = cmulf(make_float2(u0.x - u1.x, u0.y - u1.y), ca);
            z[fpad(i2)] = make_float2(u2.x + u3.x, u2.y + u3.y);
            z[fpad(i3)] = cmulf(make_float2(u2.x - u3.x, u2.y - u3.y), ca);
        }
        __syncthreads();
    }
    if (s == 0) {
        for (int b = tid; b < (L >> 1); b += 256) {
            const float2 x = z[fpad(2 * b)], y = z[fpad(2 * b + 1)];
            z[fpad(2 * b)] = make_float2(x.x + y.x, x.y + y.y);
            z[fpad(2 * b + 1)] = make_float2(x.x - y.x, x.y - y.y);
        }
        __syncthreads();
    }
}

}  // namespace rb
