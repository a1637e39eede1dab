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

// ---- convolution of two paths with radix-16 passes, L = 2^10..2^14 (FP32).  The stages run in
// passes over groups of 16 points held in registers (3-5 LDS passes and barriers per transform
// instead of 6-7 for radix 4), and the pass either side of the spectral product runs as one.
// Element i0 + k span of a group sits in slot fpad(i0) + k span + (k span) / 32 for every
// power-of-two span.  Two orders:
//  - natural (fft_convolve_r16): forward by decimation in frequency from natural order, inverse
//    by decimation in time back to natural order, the spectrum and G in bit-reversed order
//    between them.  The top stage (distance L/2) of each transform is left to the caller:
//    forward while the input is written (its upper half is the zero padding, so the stage gives
//    x_i and x_i W_L^i), inverse while the output is read (y_i = x_i + conj(W_L^i) x_{i + L/2},
//    only i < L/2 being needed).
//  - bit-reversed (fft_convolve_r16_br): forward by decimation in time from bit-reversed input to
//    natural order, G in natural order, inverse by decimation in frequency to bit-reversed output.

// Stage SB + r of a group's points (distance 2^r in the group, 2^(SB + r) in the transform),
// twiddle W = W_{2^(SB + r + 1)}^{pos + 2^SB m}, m = k mod 2^r, conjugated for the inverse (CONJ):
// decimation in time maps (a, b) to (a + W b, a - W b), decimation in frequency to
// (a + b, (a - b) W).  The upper half of a stage's twiddles is W_4 = -i (conjugate: +i) times the
// lower half.
template <int LOG2L, int SB, int R, int r, bool DIT, bool CONJ>
__host__ __device__ __forceinline__ void fft_stage(float2 (&x)[1 << R], const float2* __restrict__ tw, int pos) {
    constexpr int h = 1 << r, G = 1 << R, span = 1 << SB;
    constexpr int NL = h > 1 ? h / 2 : 1;  // twiddles read from the table
    float2 w[h];
    if constexpr (SB == 0 && r == 0) {
        w[0] = make_float2(1.f, 0.f);
    } else {
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const float2 c = tw[(pos + span * m) << (LOG2L - SB - r - 1)];
            w[m] = CONJ ? make_float2(c.x, -c.y) : c;
        }
#pragma unroll
        for (int m = NL; m < h; ++m)
            w[m] = CONJ ? make_float2(-w[m - NL].y, w[m - NL].x) : make_float2(w[m - NL].y, -w[m - NL].x);
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k & h) continue;
        const int m = k & (h - 1);
        // unit twiddles at the first two stages of a transform (pos = 0): no multiply
        const bool one = SB == 0 && m == 0 && r <= 1;
        const bool rot = SB == 0 && m == 1 && r == 1;
        const float2 a = x[k], b = x[k + h];
        if constexpr (DIT) {
            float2 y;
            if (one)
                y = b;
            else if (rot)
                y = CONJ ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
            else
                y = cmulf(w[m], b);
            x[k] = make_float2(a.x + y.x, a.y + y.y);
            x[k + h] = make_float2(a.x - y.x, a.y - y.y);
        } else {
            const float2 d = make_float2(a.x - b.x, a.y - b.y);
            x[k] = make_float2(a.x + b.x, a.y + b.y);
            if (one)
                x[k + h] = d;
            else if (rot)
                x[k + h] = CONJ ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
            else
                x[k + h] = cmulf(d, w[m]);
        }
    }
}

// stages SB .. SB + R - 1 of a group: decimation in time upwards, in frequency downwards
template <int LOG2L, int SB, int R, bool CONJ>
__host__ __device__ __forceinline__ void fft_stages_dit(float2 (&x)[1 << R], const float2* __restrict__ tw, int pos) {
    fft_stage<LOG2L, SB, R, 0, true, CONJ>(x, tw, pos);
    if constexpr (R > 1) fft_stage<LOG2L, SB, R, 1, true, CONJ>(x, tw, pos);
    if constexpr (R > 2) fft_stage<LOG2L, SB, R, 2, true, CONJ>(x, tw, pos);
    if constexpr (R > 3) fft_stage<LOG2L, SB, R, 3, true, CONJ>(x, tw, pos);
}
template <int LOG2L, int SB, int R, bool CONJ>
__host__ __device__ __forceinline__ void fft_stages_dif(float2 (&x)[1 << R], const float2* __restrict__ tw, int pos) {
    if constexpr (R > 3) fft_stage<LOG2L, SB, R, 3, false, CONJ>(x, tw, pos);
    if constexpr (R > 2) fft_stage<LOG2L, SB, R, 2, false, CONJ>(x, tw, pos);
    if constexpr (R > 1) fft_stage<LOG2L, SB, R, 1, false, CONJ>(x, tw, pos);
    fft_stage<LOG2L, SB, R, 0, false, CONJ>(x, tw, pos);
}

enum FftPassKind {
    FFT_FWD_DIF,  // forward, decimation in frequency
    FFT_INV_DIT,  // inverse, decimation in time
    FFT_MID_NAT,  // forward DIF, product with G, inverse DIT (natural order)
    FFT_FWD_DIT,  // forward, decimation in time
    FFT_INV_DIF,  // inverse, decimation in frequency
    FFT_MID_BR,   // forward DIT, product with G, inverse DIF (bit-reversed order)
};

// group t of a pass over stages SB .. SB + R - 1
template <int LOG2L, int SB, int R, int KIND>
__host__ __device__ __forceinline__ void fft_group(float2* z, const float2* __restrict__ tw,
                                                   const float2* __restrict__ G, int t) {
    constexpr int span = 1 << SB, NP = 1 << R;
    const int pos = t & (span - 1);
    const int i0 = ((t >> SB) << (SB + R)) + pos;
    const int q = fpad(i0);
    float2 x[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) x[k] = z[q + k * span + ((k * span) >> 5)];
    if constexpr (KIND == FFT_FWD_DIF || KIND == FFT_MID_NAT) fft_stages_dif<LOG2L, SB, R, false>(x, tw, pos);
    if constexpr (KIND == FFT_FWD_DIT || KIND == FFT_MID_BR) fft_stages_dit<LOG2L, SB, R, false>(x, tw, pos);
    if constexpr (KIND == FFT_MID_NAT || KIND == FFT_MID_BR) {
#pragma unroll
        for (int k = 0; k < NP; ++k) x[k] = cmulf(x[k], G[i0 + k * span]);
    }
    if constexpr (KIND == FFT_INV_DIT || KIND == FFT_MID_NAT) fft_stages_dit<LOG2L, SB, R, true>(x, tw, pos);
    if constexpr (KIND == FFT_INV_DIF || KIND == FFT_MID_BR) fft_stages_dif<LOG2L, SB, R, true>(x, tw, pos);
#pragma unroll
    for (int k = 0; k < NP; ++k) z[q + k * span + ((k * span) >> 5)] = x[k];
}

// The runner of a pass: a workgroup of 256 threads takes every group of the pass, then meets at a
// barrier (below 256 groups, L = 2^10 and 2^11, the first NG threads take one group each).  The
// schedules below are templates on the runner, so a host runner that walks the groups in order
// executes the same passes on the CPU.
struct FftWorkgroup {
    int tid;
    template <int LOG2L, int SB, int R, int KIND>
    __device__ __forceinline__ void pass(float2* z, const float2* __restrict__ tw, const float2* __restrict__ G) const {
        constexpr int NG = (1 << LOG2L) >> R;
        if constexpr (NG >= 256) {
            static_assert(NG % 256 == 0, "whole groups per thread");
#pragma unroll 2
            for (int u = 0; u < NG / 256; ++u) fft_group<LOG2L, SB, R, KIND>(z, tw, G, tid + 256 * u);
        } else {
            if (tid < NG) fft_group<LOG2L, SB, R, KIND>(z, tw, G, tid);
        }
        __syncthreads();
    }
};

// ---- natural order: stages 0 .. LOG2L - 2 in passes, the middle one at the bottom
template <int LOG2L>
constexpr int fft_mid_stages() { return (LOG2L - 1) % 4 == 0 ? 4 : (LOG2L - 1) % 4; }

template <int LOG2L, int S, class X>  // inverse passes from stage S up to LOG2L - 2
__host__ __device__ __forceinline__ void fft_nat_inv(float2* z, const float2* __restrict__ tw, const X& x) {
    if constexpr (S <= LOG2L - 2) {
        x.template pass<LOG2L, S, 4, FFT_INV_DIT>(z, tw, nullptr);
        fft_nat_inv<LOG2L, S + 4>(z, tw, x);
    }
}

template <int LOG2L, int S, class X>  // forward passes from stage S down to the middle pass
__host__ __device__ __forceinline__ void fft_nat_fwd(float2* z, const float2* __restrict__ tw,
                                                     const float2* __restrict__ G, const X& x) {
    constexpr int RM = fft_mid_stages<LOG2L>();
    if constexpr (S >= RM) {
        x.template pass<LOG2L, S - 3, 4, FFT_FWD_DIF>(z, tw, nullptr);
        fft_nat_fwd<LOG2L, S - 4>(z, tw, G, x);
    } else {
        static_assert(S == RM - 1, "passes of four stages down to the middle pass");
        x.template pass<LOG2L, 0, RM, FFT_MID_NAT>(z, tw, G);
        fft_nat_inv<LOG2L, RM>(z, tw, x);
    }
}

// Stages 0 .. LOG2L - 2 of the forward transform, the product with G (bit-reversed order) and
// stages 0 .. LOG2L - 2 of the inverse; z holds the input after the forward top stage and is left
// ready for the inverse top stage.  With FftWorkgroup: starts after and ends with a barrier.
template <int LOG2L, class X>
__host__ __device__ __forceinline__ void fft_convolve_r16(float2* z, const float2* __restrict__ tw,
                                                          const float2* __restrict__ G, const X& x) {
    static_assert(LOG2L >= 10 && LOG2L <= 14, "transform lengths 2^10 .. 2^14");
    fft_nat_fwd<LOG2L, LOG2L - 2>(z, tw, G, x);
}

// ---- bit-reversed order: forward passes from stage 0 up, the middle one at the top
template <int LOG2L, int S, class X>  // inverse passes from stage S down to 0
__host__ __device__ __forceinline__ void fft_br_inv(float2* z, const float2* __restrict__ tw, const X& x) {
    if constexpr (S >= 0) {
        x.template pass<LOG2L, S - 3, 4, FFT_INV_DIF>(z, tw, nullptr);
        fft_br_inv<LOG2L, S - 4>(z, tw, x);
    }
}

template <int LOG2L, int SB, class X>  // forward passes from stage SB up to the middle pass
__host__ __device__ __forceinline__ void fft_br_fwd(float2* z, const float2* __restrict__ tw,
                                                    const float2* __restrict__ G, const X& x) {
    constexpr int R = LOG2L - SB >= 4 ? 4 : LOG2L - SB;
    if constexpr (SB + R == LOG2L) {
        static_assert(SB % 4 == 0 && SB >= 4, "passes of four stages below the middle pass");
        x.template pass<LOG2L, SB, R, FFT_MID_BR>(z, tw, G);
        fft_br_inv<LOG2L, SB - 1>(z, tw, x);
    } else {
        x.template pass<LOG2L, SB, 4, FFT_FWD_DIT>(z, tw, nullptr);
        fft_br_fwd<LOG2L, SB + 4>(z, tw, G, x);
    }
}

// Circular convolution of z (bit-reversed order in and out, unnormalised) with the spectrum G
// (natural order).  With FftWorkgroup: starts after and ends with a barrier.
template <int LOG2L, class X>
__host__ __device__ __forceinline__ void fft_convolve_r16_br(float2* z, const float2* __restrict__ tw,
                                                             const float2* __restrict__ G, const X& x) {
    static_assert(LOG2L >= 12 && LOG2L <= 14, "radix-16 passes of at least 256 groups");
    fft_br_fwd<LOG2L, 0>(z, tw, G, x);
}

}  // namespace rb
