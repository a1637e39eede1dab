// rb_fft.hip -- Volterra convolution by FFT in LDS, for long paths (FP32 data).
//
// Replaces the length-(2N-1) FFTW convolution of updateWtilde (rBergLib/rBergomiMT.cpp:83-109)
// for N > 1024, where the Toeplitz GEMM's W1 tile no longer fits LDS.  One workgroup convolves
// two paths at once: z = W1_a + i W1_b, zero-padded to L = 2^k >= 2N-1.  Because Gamma is real,
// IFFT(FFT(z) * FFT(Gamma)) = (Gamma * W1_a) + i (Gamma * W1_b), so no separation step is
// needed.  Forward transform: decimation in time from bit-reversed input to natural order;
// pointwise product with the per-H table FFT(Gamma) (computed in FP64 on the host); inverse:
// decimation in frequency from natural order to bit-reversed output, which the epilogue reads
// back at bit-reversed indices.  L = 2^12..2^14 take radix-16 passes with the last forward pass,
// the product and the first inverse pass in one; shorter transforms radix-4 passes.  Twiddles
// exp(-2 pi i k / L) come from an FP64 table rounded to FP32.  Epilogue in FP64:
// Wtilde = sc (Y + a W1 + b W1perp).
#include <hip/hip_runtime.h>

#include "rb_fft_lds.h"
#include "rb_kernels.h"
#include "rb_types.h"

namespace rb {

// Grid: (ceil(np/2), units).  Block 256.  LDS: L + L/32 float2 (element i in slot fpad(i)).
// LOG2L: the transform length fixed at compile time (10..14), or 0 for any length.
template <int LOG2L>
__global__ void __launch_bounds__(256)
k_volterra_fft(const double* __restrict__ Z, long zstride, long np, int N, int log2L,
               const Unit* __restrict__ units, const float2* __restrict__ ghat,
               const float2* __restrict__ tw, double* __restrict__ Wt) {
    extern __shared__ float2 z[];
    const int L = 1 << log2L;
    const int u = blockIdx.y;
    const long pa = 2L * blockIdx.x, pb = pa + 1;
    const bool hasb = pb < np;
    const double* Wa = Z + pa * zstride;
    const double* Wb = Z + pb * zstride;
    // load in bit-reversed order, zero padding past N
    for (int i = threadIdx.x; i < L; i += 256) {
        const float re = i < N ? float(Wa[i]) : 0.f;
        const float im = (i < N && hasb) ? float(Wb[i]) : 0.f;
        z[fpad(int(__brev(unsigned(i)) >> (32 - log2L)))] = make_float2(re, im);
    }
    __syncthreads();
    const float2* G = ghat + long(u) * L;  // FFT(Gamma) of this unit
    if constexpr (LOG2L >= 12) {
        fft_convolve_r16_br<LOG2L>(z, tw, G, FftWorkgroup{int(threadIdx.x)});
    } else {
        if constexpr (LOG2L > 0)
            fft_dit_r4_t<LOG2L>(z, tw, threadIdx.x);
        else
            fft_dit_r4(z, tw, log2L, threadIdx.x);
        for (int k = threadIdx.x; k < L; k += 256) z[fpad(k)] = cmulf(z[fpad(k)], G[k]);
        __syncthreads();
        if constexpr (LOG2L > 0)
            fft_dif_inv_r4_t<LOG2L>(z, tw, threadIdx.x);
        else
            fft_dif_inv_r4(z, tw, log2L, threadIdx.x);
    }
    const Unit un = units[u];
    const double invL = 1.0 / double(L);
    for (int i = threadIdx.x; i < N; i += 256) {
        const float2 y = z[fpad(int(__brev(unsigned(i)) >> (32 - log2L)))];
        {
            const double Y = double(y.x) * invL;
            Wt[(long(u) * np + pa) * N + i] = un.sc * Y + un.sc * (un.a * Wa[i] + un.b * Wa[N + i]);
        }
        if (hasb) {
            const double Y = double(y.y) * invL;
            Wt[(long(u) * np + pb) * N + i] = un.sc * Y + un.sc * (un.a * Wb[i] + un.b * Wb[N + i]);
        }
    }
}

size_t volterra_fft_lds_bytes(int log2L) { return fft_lds_slots(log2L) * sizeof(float2); }

hipError_t launch_volterra_fft(const double* Z, long zstride, long np, int N, int log2L, int nunits,
                               const Unit* units, const float2* ghat, const float2* tw, double* Wt,
                               hipStream_t st) {
    const size_t lds = volterra_fft_lds_bytes(log2L);
    dim3 grid(unsigned((np + 1) / 2), unsigned(nunits));
    auto go = [&](auto kern) -> hipError_t {
        if (lds > 65536) {
            const hipError_t e = allow_lds(reinterpret_cast<const void*>(kern), lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, Z, zstride, np, N, log2L, units, ghat, tw, Wt);
        return hipGetLastError();
    };
    switch (log2L) {
        case 10: return go(&k_volterra_fft<10>);
        case 11: return go(&k_volterra_fft<11>);
        case 12: return go(&k_volterra_fft<12>);
        case 13: return go(&k_volterra_fft<13>);
        case 14: return go(&k_volterra_fft<14>);
        default: return go(&k_volterra_fft<0>);
    }
}

}  // namespace rb
