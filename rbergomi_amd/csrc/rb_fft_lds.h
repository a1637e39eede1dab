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

// ---- the same transforms with the length a compile-time constant (1024 <= L): every pass
// unrolled, butterfly spans known, and from span 32 up the four LDS slots of a butterfly at fixed
// offsets from the first (element i + k span is slot fpad(i) + k (span + span / 32)).
template <int LOG2L, int S>
__device__ __forceinline__ void fft_dit_pass(float2* z, const float2* __restrict__ tw, int tid) {
    if constexpr (S < LOG2L) {
        constexpr int L = 1 << LOG2L, span = 1 << S;
        constexpr int step = span + (span >> 5);
#pragma unroll 2
        for (int u = 0; u < (L >> 2) / 256; ++u) {
            const int t = tid + 256 * u;
            const int pos = t & (span - 1);
            const int i0 = ((t >> S) << (S + 2)) + pos;
            int q0, q1, q2, q3;
            if constexpr (span >= 32) {
                q0 = fpad(i0);
                q1 = q0 + step;
                q2 = q1 + step;
                q3 = q2 + step;
            } else {
                q0 = fpad(i0);
                q1 = fpad(i0 + span);
                q2 = fpad(i0 + 2 * span);
                q3 = fpad(i0 + 3 * span);
            }
            const float2 a = tw[pos << (LOG2L - S - 1)], b = tw[pos << (LOG2L - S - 2)];
            const float2 x0 = z[q0], x1 = z[q1], x2 = z[q2], x3 = z[q3];
            const float2 t1 = cmulf(a, x1), t3 = cmulf(a, x3);
            const float2 x0p = make_float2(x0.x + t1.x, x0.y + t1.y), x1p = make_float2(x0.x - t1.x, x0.y - t1.y);
            const float2 x2p = make_float2(x2.x + t3.x, x2.y + t3.y), x3p = make_float2(x2.x - t3.x, x2.y - t3.y);
            const float2 uu = cmulf(b, x2p), v = cmulf(b, x3p);
            z[q0] = make_float2(x0p.x + uu.x, x0p.y + uu.y);
            z[q2] = make_float2(x0p.x - uu.x, x0p.y - uu.y);
            z[q1] = make_float2(x1p.x + v.y, x1p.y - v.x);
            z[q3] = make_float2(x1p.x - v.y, x1p.y + v.x);
        }
        __syncthreads();
        fft_dit_pass<LOG2L, S + 2>(z, tw, tid);
    }
}

template <int LOG2L>
__device__ __forceinline__ void fft_dit_r4_t(float2* z, const float2* __restrict__ tw, int tid) {
    static_assert(LOG2L >= 10, "at least 256 radix-4 butterflies per pass");
    constexpr int L = 1 << LOG2L;
    if constexpr (LOG2L & 1) {
#pragma unroll 2
        for (int u = 0; u < (L >> 1) / 256; ++u) {
            const int b = tid + 256 * u;
            const int q0 = fpad(2 * b), q1 = fpad(2 * b + 1);
            const float2 x = z[q0], y = z[q1];
            z[q0] = make_float2(x.x + y.x, x.y + y.y);
            z[q1] = make_float2(x.x - y.x, x.y - y.y);
        }
        __syncthreads();
        fft_dit_pass<LOG2L, 1>(z, tw, tid);
    } else {
        fft_dit_pass<LOG2L, 0>(z, tw, tid);
    }
}

template <int LOG2L, int S>
__device__ __forceinline__ void fft_dif_pass(float2* z, const float2* __restrict__ tw, int tid) {
    constexpr int L = 1 << LOG2L;
    if constexpr (S >= 1) {
        constexpr int sh = S - 1, span1 = 1 << sh;
        constexpr int step = span1 + (span1 >> 5);
#pragma unroll 2
        for (int u = 0; u < (L >> 2) / 256; ++u) {
            const int t = tid + 256 * u;
            const int pos = t & (span1 - 1);
            const int i0 = ((t >> sh) << (S + 1)) + pos;
            int q0, q1, q2, q3;
            if constexpr (span1 >= 32) {
                q0 = fpad(i0);
                q1 = q0 + step;
                q2 = q1 + step;
                q3 = q2 + step;
            } else {
                q0 = fpad(i0);
                q1 = fpad(i0 + span1);
                q2 = fpad(i0 + 2 * span1);
                q3 = fpad(i0 + 3 * span1);
            }
            const float2 b0 = tw[pos << (LOG2L - S - 1)], a0 = tw[pos << (LOG2L - S)];
            const float2 cb = make_float2(b0.x, -b0.y), ca = make_float2(a0.x, -a0.y);
            const float2 x0 = z[q0], x1 = z[q1], x2 = z[q2], x3 = z[q3];
            const float2 u0 = make_float2(x0.x + x2.x, x0.y + x2.y);
            const float2 u2 = cmulf(make_float2(x0.x - x2.x, x0.y - x2.y), cb);
            const float2 u1 = make_float2(x1.x + x3.x, x1.y + x3.y);
            const float2 w = cmulf(make_float2(x1.x - x3.x, x1.y - x3.y), cb);
            const float2 u3 = make_float2(-w.y, w.x);
            z[q0] = make_float2(u0.x + u1.x, u0.y + u1.y);
            z[q1] = cmulf(make_float2(u0.x - u1.x, u0.y - u1.y), ca);
            z[q2] = make_float2(u2.x + u3.x, u2.y + u3.y);
            z[q3] = cmulf(make_float2(u2.x - u3.x, u2.y - u3.y), ca);
        }
        __syncthreads();
        fft_dif_pass<LOG2L, S - 2>(z, tw, tid);
    } else if constexpr (S == 0) {
#pragma unroll 2
        for (int u = 0; u < (L >> 1) / 256; ++u) {
            const int b = tid + 256 * u;
            const int q0 = fpad(2 * b), q1 = fpad(2 * b + 1);
            const float2 x = z[q0], y = z[q1];
            z[q0] = make_float2(x.x + y.x, x.y + y.y);
            z[q1] = make_float2(x.x - y.x, x.y - y.y);
        }
        __syncthreads();
    }
}

template <int LOG2L>
__device__ __forceinline__ void fft_dif_inv_r4_t(float2* z, const float2* __restrict__ tw, int tid) {
    static_assert(LOG2L >= 10, "at least 256 radix-4 butterflies per pass");
    fft_dif_pass<LOG2L, LOG2L - 1>(z, tw, tid);
}

}  // namespace rb
