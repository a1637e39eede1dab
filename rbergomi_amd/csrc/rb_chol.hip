// rb_chol.hip -- Wtilde by the exact Cholesky factor (FP64), the sampling step of RfBm::generate
// (PremiaRBergomi/RfBm.cpp:82-95) as a batched product.
//
// Per unit (H) the host factor L = [L21 | L22] (n rows of 2n, lower triangular blocks) turns the
// path's Gaussians x = (x1, x2) (streams 0 and 1 of the staged buffer, x1 doubling as W1) into
//   Wtilde[p][i] = sum_{j <= i} L21[i][j] x1[p][j] + sum_{j <= i} L22[i][j] x2[p][j].
// One workgroup computes a 64-path x 64-step tile, 4 x 4 outputs per thread, over k-chunks of
// 16 staged through LDS; chunks above the tile's last step are skipped (triangular blocks).
#include <hip/hip_runtime.h>

#include "rb_kernels.h"

namespace rb {

constexpr int CH_BM = 64, CH_BN = 64, CH_BK = 16;

// Grid: (ceil(np/64), ceil(n/64), units).  Block 256.
__global__ void __launch_bounds__(256)
k_volterra_chol_f64(const double* __restrict__ Z, long zstride, long np, int n,
                    const double* __restrict__ Lc, double* __restrict__ Wt) {
    __shared__ double Xs[CH_BK][CH_BM + 1];
    __shared__ double Ls[CH_BK][CH_BN + 1];
    const int u = blockIdx.z;
    const long p0 = long(blockIdx.x) * CH_BM;
    const int i0 = blockIdx.y * CH_BN;
    const double* L = Lc + size_t(u) * n * 2 * n;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    const int kend = min(n, i0 + CH_BN);  // steps j <= i of this tile
    // loader mapping: 64 rows x 16 k; thread -> row r = tid / 4, k = 4 (tid % 4) .. +3
    const int lr = threadIdx.x >> 2, lk = (threadIdx.x & 3) * 4;
    for (int half = 0; half < 2; ++half) {
        for (int k0 = 0; k0 < kend; k0 += CH_BK) {
            {
                const long p = p0 + lr;
                const double* x = Z + p * zstride + long(half) * n;
                const int i = i0 + lr;
                const double* l = L + size_t(i) * 2 * n + size_t(half) * n;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = k0 + lk + q;
                    Xs[lk + q][lr] = (p < np && k < n) ? x[k] : 0.0;
                    Ls[lk + q][lr] = (i < n && k < n) ? l[k] : 0.0;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < CH_BK; ++k) {
                double xa[4], lb[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) xa[a] = Xs[k][ty * 4 + a];
#pragma unroll
                for (int b = 0; b < 4; ++b) lb[b] = Ls[k][tx * 4 + b];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = fma(xa[a], lb[b], acc[a][b]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const long p = p0 + ty * 4 + a;
        if (p >= np) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + tx * 4 + b;
            if (i < n) Wt[(long(u) * np + p) * n + i] = acc[a][b];
        }
    }
}

hipError_t launch_volterra_chol(const double* Z, long zstride, long np, int n, int nunits,
                                const double* Lc, double* Wt, hipStream_t st) {
    dim3 grid(unsigned((np + CH_BM - 1) / CH_BM), unsigned((n + CH_BN - 1) / CH_BN), unsigned(nunits));
    hipLaunchKernelGGL(k_volterra_chol_f64, grid, dim3(256), 0, st, Z, zstride, np, n, Lc, Wt);
    return hipGetLastError();
}

}  // namespace rb
