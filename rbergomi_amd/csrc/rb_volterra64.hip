// rb_volterra64.hip -- the FP64 Volterra convolution on the matrix cores (N <= 1024).
//
// Same definition as k_volterra_f64 (updateWtilde, rBergLib/rBergomiMT.cpp:83-109):
//   Wtilde_i = sc (Y_i + a W1_i + b W1perp_i),  Y_i = sum_{k=1}^{i} Gamma_k W1_{i-k},
// with Y of 16 paths at a time as a lower-triangular Toeplitz product on v_mfma_f64_16x16x4_f64:
//   D[i][p] += A[i][k] B[k][p],  A[i][k] = Gamma_{i-k} (0 for i <= k),  B[k][p] = W1 of path p, step k.
// Lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds D[4 r + l/16][l%16], r = 0..3
// (checked on the device by rb_probe_mfma_layout, mode 3).  An A fragment depends only on the
// diagonal offset i0 - k0 of its 16 x 4 block, so it is one LDS read of a zero-padded copy of
// Gamma; B is one LDS read of the workgroup's W1 tile, stored k-major.  The four waves take the
// 16-row output blocks in pairs of equal cost (blocks w and 7 - w of every 8), and only the
// k-chunks at or below the diagonal run: 4 (rb + 1) MFMAs for row block rb.
// FP64 products and sums throughout; the summation order differs from the direct sum (within the
// 1e-10 relative per-path parity budget).
#include <hip/hip_runtime.h>

#include "rb_kernels.h"
#include "rb_mfma.h"
#include "rb_types.h"

namespace rb {

constexpr int V64_PATHS = 16;  // paths per workgroup (the product's N dimension)
constexpr int V64_STR = 17;    // LDS row stride of the k-major W1 tile, in doubles

size_t volterra_mfma_lds_bytes(int N) {
    const int Np = (N + 15) / 16 * 16;
    return size_t(16 + Np) * sizeof(double) + size_t(Np) * V64_STR * sizeof(double);
}

// Grid: (ceil(np / 16), units).  Block 256.  Dynamic LDS volterra_mfma_lds_bytes(N).
__global__ void __launch_bounds__(256)
k_volterra_mfma_f64(const double* __restrict__ Z, long zstride, long np, int N,
                    const Unit* __restrict__ units, const double* __restrict__ gamma,
                    double* __restrict__ Wt) {
    extern __shared__ double lds[];
    const int Np = (N + 15) & ~15;
    double* Gs = lds;            // Gs[16 + t] = Gamma_t for 1 <= t < N, 0 for t in [-16, 0] and [N, Np)
    double* Ws = lds + 16 + Np;  // Ws[k * V64_STR + p] = W1 of path p0 + p at step k (0 past N, np)
    const int u = blockIdx.y;
    const Unit un = units[u];
    const double* gu = gamma + long(u) * N;
    const long p0 = long(blockIdx.x) * V64_PATHS;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int t = tid; t < 16 + Np; t += 256) {
        const int s = t - 16;
        Gs[t] = (s >= 1 && s < N) ? gu[s] : 0.0;
    }
    // wave w loads paths w, w + 4, w + 8, w + 12, 64 consecutive steps per read; sixteen reads
    // are in flight before their LDS stores (a read-store loop would wait out each read)
    for (int k0 = 0; k0 < Np; k0 += 256) {
        double v[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const long pp = p0 + wave + 4 * a;
            const double* src = Z + pp * zstride;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + 64 * j + lane;
                v[a][j] = (pp < np && k < N) ? src[k] : 0.0;
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + 64 * j + lane;
                if (k < Np) Ws[k * V64_STR + wave + 4 * a] = v[a][j];
            }
    }
    __syncthreads();
    const int m = lane & 15, kq = lane >> 4;
    const int nb = Np >> 4;
    for (int g = 0; 8 * g < nb; ++g) {
#pragma unroll 1
        for (int half = 0; half < 2; ++half) {
            const int rb = 8 * g + (half ? 7 - wave : wave);
            if (rb >= nb) continue;
            const int i0 = 16 * rb;
            // chunk c covers steps k = 4c .. 4c + 3; rows i0 .. i0 + 15 need k <= i0 + 14, so
            // c = 0 .. 4 rb + 3: 4 (rb + 1) chunks (a multiple of 4)
            const int nc = 4 * (rb + 1);
            const double* ga = Gs + 16 + i0 + m - kq;    // A value of chunk c: ga[-4c]
            const double* wb = Ws + kq * V64_STR + m;    // B value of chunk c: wb[4c V64_STR]
            // W1perp of the epilogue, read before the products so that the reads are in flight
            // while they run
            const long p = p0 + m;
            const double* W1p = Z + p * zstride + N;
            double w1p[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 4 * r + kq;
                w1p[r] = (p < np && i < N) ? W1p[i] : 0.0;
            }
            doublex4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
            // operands of four chunks in registers; the next four are read while these multiply
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = ga[-4 * q];
                b[q] = wb[4 * q * V64_STR];
            }
            for (int c = 0; c < nc; c += 4) {
                double an[4], bn[4];
                const int cn = c + 4 < nc ? c + 4 : c;  // the last round rereads its own chunks
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    an[q] = ga[-4 * (cn + q)];
                    bn[q] = wb[4 * (cn + q) * V64_STR];
                }
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q + 1], b[q + 1], acc1, 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = an[q];
                    b[q] = bn[q];
                }
            }
            if (p >= np) continue;
            double* out = Wt + (long(u) * np + p) * N;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 4 * r + kq;
                if (i < N) {
                    const double y = acc0[r] + acc1[r];
                    out[i] = un.sc * y + un.sc * (un.a * Ws[i * V64_STR + m] + un.b * w1p[r]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- N > 1024
// The W1 tile of all steps no longer fits in LDS.  A workgroup owns 16 paths and a range of
// V64_RR output rows (32 row blocks, 8 per wave in the same equal-cost pairs), keeps their
// accumulators in registers, and streams W1 through LDS in segments of S steps (S from the LDS
// left after the zero-padded Gamma).  A row block takes, in each segment, the chunks of that
// segment at or below its diagonal.  Consecutive workgroups share their paths (row range
// fastest), so the repeated W1 reads of one path tile meet in L2.
constexpr int V64_RR = 512;

size_t volterra_mfma_long_lds_bytes(int N, int S) {
    const int Np = (N + 15) / 16 * 16;
    return size_t(16 + Np) * sizeof(double) + size_t(S) * V64_STR * sizeof(double);
}

// Grid: (row ranges x path tiles, units), x = rr + nrr * tile.  Block 256.
__global__ void __launch_bounds__(256)
k_volterra_mfma_long_f64(const double* __restrict__ Z, long zstride, long np, int N, int S,
                         const Unit* __restrict__ units, const double* __restrict__ gamma,
                         double* __restrict__ Wt) {
    extern __shared__ double lds[];
    const int Np = (N + 15) & ~15;
    double* Gs = lds;            // as in k_volterra_mfma_f64
    double* Ws = lds + 16 + Np;  // Ws[kk * V64_STR + p] = W1 of path p0 + p at step ks + kk
    const int nrr = (Np + V64_RR - 1) / V64_RR;
    const int rr = int(blockIdx.x % unsigned(nrr));
    const long p0 = long(blockIdx.x / unsigned(nrr)) * V64_PATHS;
    const int u = blockIdx.y;
    const Unit un = units[u];
    const double* gu = gamma + long(u) * N;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int R0 = rr * V64_RR, R1 = R0 + V64_RR < Np ? R0 + V64_RR : Np;
    for (int t = tid; t < 16 + Np; t += 256) {
        const int s = t - 16;
        Gs[t] = (s >= 1 && s < N) ? gu[s] : 0.0;
    }
    // row block h of this wave: local block 8 (h / 2) + (h odd ? 7 - wave : wave)
    doublex4_t acc0[8], acc1[8];
#pragma unroll
    for (int h = 0; h < 8; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc0[h][r] = acc1[h][r] = 0.0;
    // steps k <= R1 - 2 feed rows below R1
    for (int ks = 0; ks < R1 - 1; ks += S) {
        __syncthreads();  // the previous segment's reads are done
        for (int k0 = 0; k0 < S; k0 += 256) {
            double v[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const long pp = p0 + wave + 4 * a;
                const double* src = Z + pp * zstride;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int kk = k0 + 64 * j + lane, k = ks + kk;
                    v[a][j] = (kk < S && pp < np && k < N) ? src[k] : 0.0;
                }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int kk = k0 + 64 * j + lane;
                    if (kk < S) Ws[kk * V64_STR + wave + 4 * a] = v[a][j];
                }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int rb = (R0 >> 4) + 8 * (h >> 1) + ((h & 1) ? 7 - wave : wave);
            if (16 * rb >= R1) continue;
            const int i0 = 16 * rb;
            // chunks c (steps 4c .. 4c + 3) of this segment at or below the diagonal: c < 4 rb + 4
            const int c_lo = ks >> 2;
            const int c_hi = min((ks + S) >> 2, 4 * rb + 4);
            const double* ga = Gs + 16 + i0 + m - kq;
            const double* wb = Ws + (kq - ks) * V64_STR + m;  // B value of chunk c: wb[4c V64_STR]
            for (int c = c_lo; c < c_hi; c += 4) {
                double a[4], b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = ga[-4 * (c + q)];
                    b[q] = wb[4 * (c + q) * V64_STR];
                }
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    acc0[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc0[h], 0, 0, 0);
                    acc1[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q + 1], b[q + 1], acc1[h], 0, 0, 0);
                }
            }
        }
    }
    const long p = p0 + m;
    if (p >= np) return;
    const double* W1 = Z + p * zstride;
    const double* W1p = W1 + N;
    double* out = Wt + (long(u) * np + p) * N;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int rb = (R0 >> 4) + 8 * (h >> 1) + ((h & 1) ? 7 - wave : wave);
        if (16 * rb >= R1) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * rb + 4 * r + kq;
            if (i < N) {
                const double y = acc0[h][r] + acc1[h][r];
                out[i] = un.sc * y + un.sc * (un.a * W1[i] + un.b * W1p[i]);
            }
        }
    }
}

// segment length for N: the LDS left after Gamma, in multiples of 16 steps, at most Np
int volterra_mfma_long_segment(int N, size_t lds_limit) {
    const int Np = (N + 15) / 16 * 16;
    const long avail = long(lds_limit) - long(16 + Np) * long(sizeof(double));
    long S = avail / long(V64_STR * sizeof(double));
    S = S / 16 * 16;
    return int(S < Np ? S : Np);
}

hipError_t launch_volterra_mfma_long_f64(const double* Z, long zstride, long np, int N, int nunits,
                                         const Unit* units, const double* gamma, double* Wt,
                                         size_t lds_limit, hipStream_t st) {
    const int S = volterra_mfma_long_segment(N, lds_limit);
    if (S < 16) return hipErrorInvalidValue;
    const size_t lds = volterra_mfma_long_lds_bytes(N, S);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_volterra_mfma_long_f64), lds);
        if (e != hipSuccess) return e;
    }
    const int Np = (N + 15) / 16 * 16;
    const long nrr = (Np + V64_RR - 1) / V64_RR;
    const long tiles = (np + V64_PATHS - 1) / V64_PATHS;
    dim3 grid(unsigned(nrr * tiles), unsigned(nunits));
    hipLaunchKernelGGL(k_volterra_mfma_long_f64, grid, dim3(256), lds, st, Z, zstride, np, N, S, units, gamma,
                       Wt);
    return hipGetLastError();
}

hipError_t launch_volterra_mfma_f64(const double* Z, long zstride, long np, int N, int nunits,
                                    const Unit* units, const double* gamma, double* Wt, hipStream_t st) {
    const size_t lds = volterra_mfma_lds_bytes(N);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_volterra_mfma_f64), lds);
        if (e != hipSuccess) return e;
    }
    dim3 grid(unsigned((np + V64_PATHS - 1) / V64_PATHS), unsigned(nunits));
    hipLaunchKernelGGL(k_volterra_mfma_f64, grid, dim3(256), lds, st, Z, zstride, np, N, units, gamma, Wt);
    return hipGetLastError();
}

}  // namespace rb
