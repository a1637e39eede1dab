// rb_fft_lds.h -- complex FP32 transforms of a power-of-two length held in LDS, shared by the
// long-path kernels (rb_fft.hip, rb_fused_fft.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rb {

// LDS slot of element i: one pad slot after every 32 elements, so the power-of-two strides of
// the butterflies and of the bit-reversed loads spread over the banks
__host__ __device__ __forceinline__ int fpad(int i) { return i + (i >> 5); }
// LDS slots for L elements
inline size_t fft_lds_slots(int log2L) { return (size_t(1) << log2L) + (size_t(1) << log2L >> 5); }

__host__ __device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
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

// ---- radix 2^R passes (R <= 4): the 2^R points of a butterfly group stay in registers for R
// radix-2 stages, so a transform of L = 2^12..2^14 takes 3-4 LDS passes and barriers instead of
// 6-7.  Element i0 + k span of a group sits in slot fpad(i0) + k span + (k span) / 32 for every
// power-of-two span (a group never straddles a pad in a way that breaks this).

// Forward group t of stages S .. S + R - 1 (decimation in time): stage S + r pairs points k and
// k + 2^r of the group (bit r of k clear) with the twiddle W_{2 span 2^r}^{pos + span m},
// m = k mod 2^r; the upper half of those twiddles is -i times the lower half.
template <int LOG2L, int S, int R>
__host__ __device__ __forceinline__ void fft_dit_group(float2* z, const float2* __restrict__ tw, int t) {
    constexpr int span = 1 << S, G = 1 << R;
    const int pos = t & (span - 1);
    const int q = fpad(((t >> S) << (S + R)) + pos);
    float2 x[G];
#pragma unroll
    for (int k = 0; k < G; ++k) x[k] = z[q + k * span + ((k * span) >> 5)];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int h = 1 << r;
        float2 w[8];
        if (S == 0 && r <= 1) {  // W_2^0 = 1, W_4^1 = -i
            w[0] = make_float2(1.f, 0.f);
            w[1] = make_float2(0.f, -1.f);
        } else {
#pragma unroll
            for (int m = 0; m < (h > 1 ? h / 2 : 1); ++m) w[m] = tw[(pos + span * m) << (LOG2L - S - r - 1)];
#pragma unroll
            for (int m = h / 2; m < h && h > 1; ++m) w[m] = make_float2(w[m - h / 2].y, -w[m - h / 2].x);
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k & h) continue;
            const int m = k & (h - 1);
            float2 y;
            if (S == 0 && r == 0)
                y = x[k + h];
            else if (S == 0 && r == 1 && m == 1)
                y = make_float2(x[k + h].y, -x[k + h].x);
            else
                y = cmulf(w[m], x[k + h]);
            x[k + h] = make_float2(x[k].x - y.x, x[k].y - y.y);
            x[k] = make_float2(x[k].x + y.x, x[k].y + y.y);
        }
    }
#pragma unroll
    for (int k = 0; k < G; ++k) z[q + k * span + ((k * span) >> 5)] = x[k];
}

// Inverse group t of stages S down to S - R + 1 (decimation in frequency, conjugate twiddles):
// stage sb + r (sb = S - R + 1, distance span1 2^r, span1 = 2^sb) maps points (k, k + 2^r) to
// (a + b, (a - b) conj W_{2 span1 2^r}^{pos + span1 m}); the upper half of those twiddles is i
// times the lower half.
template <int LOG2L, int S, int R>
__host__ __device__ __forceinline__ void fft_dif_group(float2* z, const float2* __restrict__ tw, int t) {
    constexpr int sb = S - R + 1, span1 = 1 << sb, G = 1 << R;
    const int pos = t & (span1 - 1);
    const int q = fpad(((t >> sb) << (S + 1)) + pos);
    float2 x[G];
#pragma unroll
    for (int k = 0; k < G; ++k) x[k] = z[q + k * span1 + ((k * span1) >> 5)];
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        const int h = 1 << r;
        float2 w[8];
        if (sb == 0 && r <= 1) {  // conj W_2^0 = 1, conj W_4^1 = i
            w[0] = make_float2(1.f, 0.f);
            w[1] = make_float2(0.f, 1.f);
        } else {
#pragma unroll
            for (int m = 0; m < (h > 1 ? h / 2 : 1); ++m) {
                const float2 c = tw[(pos + span1 * m) << (LOG2L - sb - r - 1)];
                w[m] = make_float2(c.x, -c.y);
            }
#pragma unroll
            for (int m = h / 2; m < h && h > 1; ++m) w[m] = make_float2(-w[m - h / 2].y, w[m - h / 2].x);
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k & h) continue;
            const int m = k & (h - 1);
            const float2 a = x[k], b = x[k + h];
            const float2 d = make_float2(a.x - b.x, a.y - b.y);
            x[k] = make_float2(a.x + b.x, a.y + b.y);
            if (sb == 0 && r == 0)
                x[k + h] = d;
            else if (sb == 0 && r == 1 && m == 1)
                x[k + h] = make_float2(-d.y, d.x);
            else
                x[k + h] = cmulf(d, w[m]);
        }
    }
#pragma unroll
    for (int k = 0; k < G; ++k) z[q + k * span1 + ((k * span1) >> 5)] = x[k];
}

// Stages S .. LOG2L - 1 of the forward transform: radix-16 passes, the remainder (radix 2, 4 or
// 8) last, where the spans are largest
template <int LOG2L, int S>
__device__ __forceinline__ void fft_dit16_pass(float2* z, const float2* __restrict__ tw, int tid) {
    if constexpr (S < LOG2L) {
        constexpr int R = LOG2L - S >= 4 ? 4 : LOG2L - S;
        constexpr int NG = (1 << LOG2L) >> R;
        static_assert(NG % 256 == 0, "whole groups per thread");
#pragma unroll 2
        for (int u = 0; u < NG / 256; ++u) fft_dit_group<LOG2L, S, R>(z, tw, tid + 256 * u);
        __syncthreads();
        fft_dit16_pass<LOG2L, S + R>(z, tw, tid);
    }
}

// Stages S down to 0 of the inverse transform: the remainder first, then radix-16 passes
template <int LOG2L, int S>
__device__ __forceinline__ void fft_dif16_pass(float2* z, const float2* __restrict__ tw, int tid) {
    if constexpr (S >= 0) {
        constexpr int R = (S + 1) % 4 == 0 ? 4 : (S + 1) % 4;
        constexpr int NG = (1 << LOG2L) >> R;
        static_assert(NG % 256 == 0, "whole groups per thread");
#pragma unroll 2
        for (int u = 0; u < NG / 256; ++u) fft_dif_group<LOG2L, S, R>(z, tw, tid + 256 * u);
        __syncthreads();
        fft_dif16_pass<LOG2L, S - R>(z, tw, tid);
    }
}

// Middle group t of a convolution: forward stages S .. S + R - 1 (the last forward pass), the
// product with G at the natural indices the group then holds, and inverse stages S + R - 1 .. S
// (the first inverse pass, which has the same groups): two LDS passes and two barriers fewer.
template <int LOG2L, int S, int R>
__host__ __device__ __forceinline__ void fft_mid_group(float2* z, const float2* __restrict__ tw,
                                                       const float2* __restrict__ G, int t) {
    constexpr int span = 1 << S, NP = 1 << R;
    const int pos = t & (span - 1);
    const int i0 = ((t >> S) << (S + R)) + pos;
    const int q = fpad(i0);
    float2 x[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) x[k] = z[q + k * span + ((k * span) >> 5)];
#pragma unroll
    for (int r = 0; r < R; ++r) {  // forward, as fft_dit_group (S >= 4 here: no unit twiddles)
        const int h = 1 << r;
        float2 w[8];
#pragma unroll
        for (int m = 0; m < (h > 1 ? h / 2 : 1); ++m) w[m] = tw[(pos + span * m) << (LOG2L - S - r - 1)];
#pragma unroll
        for (int m = h / 2; m < h && h > 1; ++m) w[m] = make_float2(w[m - h / 2].y, -w[m - h / 2].x);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            if (k & h) continue;
            const float2 y = cmulf(w[k & (h - 1)], x[k + h]);
            x[k + h] = make_float2(x[k].x - y.x, x[k].y - y.y);
            x[k] = make_float2(x[k].x + y.x, x[k].y + y.y);
        }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) x[k] = cmulf(x[k], G[i0 + k * span]);
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {  // inverse, as fft_dif_group with sb = S
        const int h = 1 << r;
        float2 w[8];
#pragma unroll
        for (int m = 0; m < (h > 1 ? h / 2 : 1); ++m) {
            const float2 c = tw[(pos + span * m) << (LOG2L - S - r - 1)];
            w[m] = make_float2(c.x, -c.y);
        }
#pragma unroll
        for (int m = h / 2; m < h && h > 1; ++m) w[m] = make_float2(-w[m - h / 2].y, w[m - h / 2].x);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            if (k & h) continue;
            const float2 a = x[k], b = x[k + h];
            x[k] = make_float2(a.x + b.x, a.y + b.y);
            x[k + h] = cmulf(make_float2(a.x - b.x, a.y - b.y), w[k & (h - 1)]);
        }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) z[q + k * span + ((k * span) >> 5)] = x[k];
}

// Circular convolution with the transform G of a real kernel: z (bit-reversed order) ->
// forward radix-16 passes -> product with G -> inverse passes -> z (bit-reversed order,
// unnormalised); the last forward and first inverse passes run as one (fft_mid_group)
template <int LOG2L, int S>
__device__ __forceinline__ void fft_conv16_pass(float2* z, const float2* __restrict__ tw,
                                                const float2* __restrict__ G, int tid) {
    constexpr int R = LOG2L - S >= 4 ? 4 : LOG2L - S;
    constexpr int NG = (1 << LOG2L) >> R;
    static_assert(NG % 256 == 0, "whole groups per thread");
    if constexpr (S + R == LOG2L) {
        static_assert(S % 4 == 0 && S >= 4, "radix-16 passes below the middle one");
#pragma unroll 2
        for (int u = 0; u < NG / 256; ++u) fft_mid_group<LOG2L, S, R>(z, tw, G, tid + 256 * u);
        __syncthreads();
        fft_dif16_pass<LOG2L, S - 1>(z, tw, tid);
    } else {
#pragma unroll 2
        for (int u = 0; u < NG / 256; ++u) fft_dit_group<LOG2L, S, R>(z, tw, tid + 256 * u);
        __syncthreads();
        fft_conv16_pass<LOG2L, S + R>(z, tw, G, tid);
    }
}

template <int LOG2L>
__device__ __forceinline__ void fft_convolve_r16_t(float2* z, const float2* __restrict__ tw,
                                                   const float2* __restrict__ G, int tid) {
    static_assert(LOG2L >= 12, "at least 256 radix-16 groups per pass");
    fft_conv16_pass<LOG2L, 0>(z, tw, G, tid);
}

// The transforms of fft_dit_r4_t / fft_dif_inv_r4_t with radix-16 passes (L >= 2^12: at least
// one 16-point group per thread)
template <int LOG2L>
__device__ __forceinline__ void fft_dit_r16_t(float2* z, const float2* __restrict__ tw, int tid) {
    static_assert(LOG2L >= 12, "at least 256 radix-16 groups per pass");
    fft_dit16_pass<LOG2L, 0>(z, tw, tid);
}

template <int LOG2L>
__device__ __forceinline__ void fft_dif_inv_r16_t(float2* z, const float2* __restrict__ tw, int tid) {
    static_assert(LOG2L >= 12, "at least 256 radix-16 groups per pass");
    fft_dif16_pass<LOG2L, LOG2L - 1>(z, tw, tid);
}

}  // namespace rb
