// rb_staged.hip -- the staged FP64 path: every stage writes its result to HBM.
//
// This is the plain FP64 single-request GPU path (parity mode: per-path results match the
// CPU restatement of the reference to ~1e-13).  Stages, with the reference code each replaces:
//   k_normals_f64    Philox Gaussians          <- genGaussianMT (rBergLib/rBergomiMT.cpp:169-172)
//   k_volterra_f64   Wtilde = sc (Gamma*W1 + a W1 + b W1perp)
//                                              <- updateWtilde (:83-109), FFTW replaced by a
//                                                 register-windowed direct Toeplitz sum
//   k_scan_rt_f64    Sum v, Sum sqrt(v) W1 per (path, group)
//                                              <- scaleWtilde/updateV/intVdt/intRootVdW (:73-80,163-167,427-436)
//   k_scan_mc_f64    log S per (path, group)   <- updateZ/scaleZ/updateS (:111-136)
//   k_payoff_f64     payoff per (path, option), block partial sums in FP64
//                                              <- BS_call_price / posPart (:383-391, :264-266)
//   k_reduce_f64     ordered sum of block partials -> price, stat   (:404-409)
//   k_iv_f64         implied vols                                    (IV_call, :421-423)
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "rb_device.h"
#include "rb_kernels.h"
#include "rb_types.h"

namespace rb {

// Raise a kernel's dynamic-LDS limit once per device (later calls with the same or a smaller size
// make no runtime call, so launches can be recorded into a graph after the first eager run).
hipError_t allow_lds(const void* fn, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = done[{dev, fn}];
    if (have >= bytes) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    if (e == hipSuccess) have = bytes;
    return e;
}

// ---------------------------------------------------------------- Gaussians
// One thread per (path, stream, 4-step block).  out[(p*nstreams + s)*N + j]
__global__ void k_normals_f64(PhiloxKey key, long p0, long np, int nstreams, int N,
                              double* __restrict__ out) {
    const int nq = (N + 3) / 4;
    const long total = np * nstreams * nq;
    for (long t = blockIdx.x * long(blockDim.x) + threadIdx.x; t < total;
         t += long(gridDim.x) * blockDim.x) {
        const int q = int(t % nq);
        const long ps = t / nq;
        const int s = int(ps % nstreams);
        const long p = ps / nstreams;
        double z[4];
        normals4_f64(key, uint64_t(p0 + p), uint32_t(s), uint32_t(q), z);
        double* o = out + ps * N + 4 * q;
        if (4 * q + 3 < N) {
            o[0] = z[0]; o[1] = z[1]; o[2] = z[2]; o[3] = z[3];
        } else {
            for (int j = 0; j < 4 && 4 * q + j < N; ++j) o[j] = z[j];
        }
    }
}

// ---------------------------------------------------------------- Brownian-bridge mapping
// hierarchical2increments (rBergomi/aux.h:105-125) on every row of N = 2^m values, in place: the
// coarse-to-fine coefficients Z of a Brownian bridge on 2^m cells become i.i.d. N(0,1) increments
//   dW_i = 2^(-m/2) Z_0 + sum_{n=1}^{m} 2^((n-m-1)/2) s_n(i) Z_{2^(n-1) + floor(k_n / 2)},
// k_n = floor(i 2^(n-m)), s_n(i) = +1 for even k_n (left half of its level-n cell), -1 for odd.
// One workgroup per row; the row is staged in LDS.
__global__ void __launch_bounds__(256) k_bridge_f64(double* __restrict__ Z, long rows, int m) {
    extern __shared__ double zr[];
    const int N = 1 << m;
    for (long r = blockIdx.x; r < rows; r += gridDim.x) {
        double* row = Z + r * N;
        for (int i = threadIdx.x; i < N; i += 256) zr[i] = row[i];
        __syncthreads();
        for (int i = threadIdx.x; i < N; i += 256) {
            double v = exp2(-0.5 * m) * zr[0];
            for (int n = 1; n <= m; ++n) {
                const int k = i >> (m - n);
                const double w = exp2(0.5 * (n - m - 1));
                v += (k & 1) ? -w * zr[(1 << (n - 1)) + (k >> 1)] : w * zr[(1 << (n - 1)) + (k >> 1)];
            }
            row[i] = v;
        }
        __syncthreads();
    }
}

hipError_t launch_bridge_f64(double* Z, long rows, int N, hipStream_t st) {
    int m = 0;
    while ((1 << m) < N) ++m;
    if ((1 << m) != N || rows < 1) return hipErrorInvalidValue;
    const size_t lds = size_t(N) * sizeof(double);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_bridge_f64), lds);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = unsigned(rows < 65536 ? rows : 65536);
    hipLaunchKernelGGL(k_bridge_f64, dim3(grid), dim3(256), lds, st, Z, rows, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------- Volterra convolution
// Y_i = sum_{k=1}^{i} Gamma_k W1_{i-k}.  A thread owns four consecutive outputs i = 4b+r and
// slides a four-entry window of W1 down the Toeplitz diagonal: per k one broadcast read of
// Gamma_k, one new W1 read and four FMAs.  W1 sits in LDS split by residue (i mod 4) so that
// the per-k read W1[4b-k-1] of consecutive threads b hits consecutive addresses.
// Block: 256 threads; paths per block ppb = 256 / ceil(N/4) (>= 1); for N > 1024 each thread
// loops over several 4-blocks of its path.  Grid: (path tiles, units).
__global__ void __launch_bounds__(256)
k_volterra_f64(const double* __restrict__ Z, long zstride, long np, int N,
               const Unit* __restrict__ units, const double* __restrict__ gamma,
               double* __restrict__ Wt) {
    extern __shared__ double lds[];
    const int nb = (N + 3) / 4;                 // 4-blocks per path
    const int tpp = nb < 256 ? nb : 256;        // threads per path
    const int ppb = 256 / tpp;                  // paths per block
    const int Q = nb + 1;                       // row length of one residue class
    double* g = lds;                            // Gamma[N]
    double* w = lds + N;                        // [ppb][4][Q]
    const int u = blockIdx.y;
    const Unit un = units[u];
    const double* gu = gamma + long(u) * N;
    for (int i = threadIdx.x; i < N; i += 256) g[i] = gu[i];
    const long pbase = long(blockIdx.x) * ppb;
    // residue layout w[slot][r][q] = W1[4q + r]; entries past N (and absent paths) are zero
    for (int t = threadIdx.x; t < ppb * 4 * Q; t += 256) {
        const int sl = t / (4 * Q), rem = t % (4 * Q);
        const int i = 4 * (rem % Q) + rem / Q;
        const long p = pbase + sl;
        w[t] = (p < np && i < N) ? Z[p * zstride + i] : 0.0;
    }
    __syncthreads();
    const int slot = threadIdx.x / tpp;
    const long p = pbase + slot;
    if (slot >= ppb || p >= np) return;
    const double* ws = w + slot * 4 * Q;
    const double* W1 = Z + p * zstride;
    const double* W1p = W1 + N;
    double* out = Wt + (long(u) * np + p) * N;
    for (int b = threadIdx.x % tpp; b < nb; b += tpp) {
        const int i0 = 4 * b;
        // window x_r = W1[i0 + r - k] for the current k, starting at k = 1
        double x0 = i0 - 1 >= 0 ? ws[3 * Q + ((i0 - 1) >> 2)] : 0.0;
        double x1 = ws[0 * Q + (i0 >> 2)];
        double x2 = ws[1 * Q + (i0 >> 2)];
        double x3 = ws[2 * Q + (i0 >> 2)];
        double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
        const int kmax = i0 + 3 < N - 1 ? i0 + 3 : N - 1;
        for (int k = 1; k <= kmax; ++k) {
            const double gk = g[k];
            y0 = fma(gk, x0, y0);
            y1 = fma(gk, x1, y1);
            y2 = fma(gk, x2, y2);
            y3 = fma(gk, x3, y3);
            // shift: x_r(k+1) = x_{r-1}(k), x_0(k+1) = W1[i0 - k - 1]
            x3 = x2; x2 = x1; x1 = x0;
            const int j = i0 - k - 1;
            x0 = j >= 0 ? ws[(j & 3) * Q + (j >> 2)] : 0.0;
        }
        const double y[4] = {y0, y1, y2, y3};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = i0 + r;
            if (i < N) out[i] = un.sc * y[r] + un.sc * (un.a * W1[i] + un.b * W1p[i]);
        }
    }
}

// ---------------------------------------------------------------- integrals (RT)
// One wave per (path, group); lanes stride over steps.  Fixed butterfly order => deterministic.
// out[(g*np + p)*2 + {0,1}] = (Ivdt, IsvdW)
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(256)
k_scan_rt_f64(const double* __restrict__ Z, long zstride, const double* __restrict__ Wt,
              long np, int N, const Group* __restrict__ groups, int ngroups,
              const double* __restrict__ comp, double* __restrict__ out, double sign) {
    const long wid = (long(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (wid >= np * ngroups) return;
    const int gi = int(wid % ngroups);  // the groups of one path are adjacent waves
    const long p = wid / ngroups;
    const Group G = groups[gi];
    const double* wt = Wt + (long(G.unit) * np + p) * N;
    const double* W1 = Z + p * zstride;
    const double* c = comp + long(gi) * N;
    const double sxi = sqrt(G.xi);
    const double e2 = 0.5 * G.eta * G.TH;
    double sv = 0.0, ss = 0.0;
    for (int i = lane; i < N; i += 64) {
        // sqrt(v_i) = sqrt(xi) exp((eta T^H Wtilde_{i-1} - c_i)/2), v_0 = xi
        // sign = -1: the antithetic path (-W1, -W1perp), whose Volterra process is -Wtilde
        // exactly (the convolution is linear and negation is exact)
        const double r = i == 0 ? sxi : sxi * exp(e2 * (sign * wt[i - 1]) - 0.5 * c[i]);
        sv = fma(r, r, sv);
        ss = fma(r, sign * W1[i], ss);
    }
    sv = wave_sum_f64(sv);
    ss = wave_sum_f64(ss);
    if (lane == 0) {
        out[(long(gi) * np + p) * 2 + 0] = G.dt * sv;
        out[(long(gi) * np + p) * 2 + 1] = G.sdt * ss;
    }
}

// ---------------------------------------------------------------- log-spot (plain MC)
// log S = sum_i sqrt(v_i) sdt (rho W1_i + sqrt(1-rho^2) Wperp_i) - v_i dt / 2
__global__ void __launch_bounds__(256)
k_scan_mc_f64(const double* __restrict__ Z, long zstride, const double* __restrict__ Wt,
              long np, int N, const Group* __restrict__ groups, int ngroups,
              const double* __restrict__ comp, double* __restrict__ out) {
    const long wid = (long(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (wid >= np * ngroups) return;
    const int gi = int(wid % ngroups);  // the groups of one path are adjacent waves
    const long p = wid / ngroups;
    const Group G = groups[gi];
    const double* wt = Wt + (long(G.unit) * np + p) * N;
    const double* W1 = Z + p * zstride;
    const double* Wp = W1 + 2 * N;
    const double* c = comp + long(gi) * N;
    const double sxi = sqrt(G.xi);
    const double e2 = 0.5 * G.eta * G.TH;
    double x = 0.0;
    for (int i = lane; i < N; i += 64) {
        const double r = i == 0 ? sxi : sxi * exp(e2 * wt[i - 1] - 0.5 * c[i]);
        const double z = G.rho * W1[i] + G.crho * Wp[i];
        x += r * (G.sdt * z) - 0.5 * (r * r) * G.dt;
    }
    x = wave_sum_f64(x);
    if (lane == 0) {
        out[(long(gi) * np + p) * 2 + 0] = x;
        out[(long(gi) * np + p) * 2 + 1] = 0.0;
    }
}

// ---------------------------------------------------------------- payoffs
// Grid (path tiles of 256, options).  Block sums (payoff, payoff^2) in FP64 with a fixed tree;
// partial[(I*nblk_total + blk0 + blockIdx.x)*2 + {0,1}].  If payoff != nullptr the per-path
// payoffs are also written at payoff[(p0+p)*out_size + I] (quadrature layout).
__global__ void __launch_bounds__(256)
k_payoff_f64(const double* __restrict__ ints, long np, long p0, const Option* __restrict__ opts,
             int out_size, int method, long blk0, long nblk_total, double* __restrict__ partial,
             double* __restrict__ payoff) {
    __shared__ double s1[256], s2[256];
    const int I = blockIdx.y;
    const Option o = opts[I];
    const long p = long(blockIdx.x) * 256 + threadIdx.x;
    double x = 0.0;
    if (p < np) {
        const double* iv = ints + (long(o.group) * np + p) * 2;
        if (method == METHOD_RT) {
            x = rt_payoff_f64(iv[0], iv[1], o.rho, o.K, o.lnK);
        } else {
            const double S = exp(iv[0]);
            x = S - o.K > 0.0 ? S - o.K : 0.0;
        }
        if (payoff) payoff[(p0 + p) * out_size + I] = x;
    }
    s1[threadIdx.x] = x;
    s2[threadIdx.x] = x * x;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && partial) {
        double* d = partial + (long(I) * nblk_total + blk0 + blockIdx.x) * 2;
        d[0] = s1[0];
        d[1] = s2[0];
    }
}

// One block per option: ordered strided sums then a fixed tree => deterministic.
// price = S/M, stat = sqrt(S2/M - price^2)/sqrt(M)   (rBergLib/rBergomiMT.cpp:404-409)
__global__ void __launch_bounds__(256)
k_reduce_f64(const double* __restrict__ partial, long nblk, int stride_blk, long M,
             double* __restrict__ price, double* __restrict__ stat) {
    __shared__ double s1[256], s2[256];
    const int I = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (long k = threadIdx.x; k < nblk; k += 256) {
        const double* d = partial + (long(I) * nblk + k) * stride_blk;
        a += d[0];
        b += d[1];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = double(M);
        const double pr = s1[0] / m;
        price[I] = pr;
        stat[I] = sqrt(s2[0] / m - pr * pr) / sqrt(m);
    }
}

// One block per output row: rows of (sum, sum of squares) pairs at partial + row_off[b],
// row_n[b] pairs long; same ordered strided sums and fixed tree as k_reduce_f64.
__global__ void __launch_bounds__(256)
k_reduce_rows(const double* __restrict__ partial, const long* __restrict__ row_off,
              const long* __restrict__ row_n, const long* __restrict__ row_M,
              double* __restrict__ price, double* __restrict__ stat) {
    __shared__ double s1[256], s2[256];
    const int I = blockIdx.x;
    const double* base = partial + row_off[I];
    const long n = row_n[I];
    double a = 0.0, b = 0.0;
    for (long k = threadIdx.x; k < n; k += 256) {
        a += base[2 * k];
        b += base[2 * k + 1];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = double(row_M[I]);
        const double pr = s1[0] / m;
        price[I] = pr;
        stat[I] = sqrt(s2[0] / m - pr * pr) / sqrt(m);
    }
}

// k_reduce_rows and k_iv_f64 in one launch (the same sums, the same vols), and the fused
// kernel's item counter set back to zero for the batch's next launch
__global__ void __launch_bounds__(256)
k_reduce_rows_iv(const double* __restrict__ partial, const long* __restrict__ row_off,
                 const long* __restrict__ row_n, const long* __restrict__ row_M,
                 const Option* __restrict__ opts, double* __restrict__ price, double* __restrict__ stat,
                 double* __restrict__ iv, int* __restrict__ counter) {
    __shared__ double s1[256], s2[256];
    const int I = blockIdx.x;
    const double* base = partial + row_off[I];
    const long n = row_n[I];
    double a = 0.0, b = 0.0;
    for (long k = threadIdx.x; k < n; k += 256) {
        a += base[2 * k];
        b += base[2 * k + 1];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = double(row_M[I]);
        const double pr = s1[0] / m;
        price[I] = pr;
        stat[I] = sqrt(s2[0] / m - pr * pr) / sqrt(m);
        iv[I] = implied_vol_f64(pr, opts[I].K, opts[I].T);
        if (I == 0) *counter = 0;
    }
}

__global__ void k_iv_f64(const double* __restrict__ price, const Option* __restrict__ opts,
                         int n, double* __restrict__ iv) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) iv[I] = implied_vol_f64(price[I], opts[I].K, opts[I].T);
}

// implied vols to the reference's tolerance (implied_vol_reference_f64)
__global__ void k_iv_ref_f64(const double* __restrict__ price, const double* __restrict__ K,
                             const double* __restrict__ T, int n, double* __restrict__ iv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iv[i] = implied_vol_reference_f64(price[i], K[i], T[i]);
}

// Philox known-answer probe: out[4*t..] = Philox4x32-10(ctr[t], key)
__global__ void k_philox_probe(const uint4* __restrict__ ctr, PhiloxKey key, long n,
                               uint4* __restrict__ out) {
    const long t = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (t < n) out[t] = philox10(ctr[t], key);
}

// ---------------------------------------------------------------- launchers
static inline int grid_for(long n, int block, int cap = 65535 * 8) {
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return int(g < cap ? g : cap);
}

hipError_t launch_normals_f64(PhiloxKey key, long p0, long np, int nstreams, int N, double* out,
                              hipStream_t st) {
    const long total = np * nstreams * ((N + 3) / 4);
    hipLaunchKernelGGL(k_normals_f64, dim3(grid_for(total, 256)), dim3(256), 0, st, key, p0, np,
                       nstreams, N, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- Sobol' Gaussians
// Point i = p0 + p + 1 of the Sobol' sequence (dimension d: XOR of the direction integers V[d][k]
// over the set bits k of i, over 2^32), mapped by the inverse normal CDF.  The point's dimensions
// are dealt to the streams in turn, dimension nstreams*j + s -> stream s, step j (for two streams:
// even dimensions to W1 and odd ones to W1perp, as breakZ does, rBergomi/aux.h:186-191).  One
// thread per (path, dimension).
// With a digital shift (shift != nullptr, one 32-bit word per dimension XORed into every point,
// which keeps the net structure and makes each point uniform), point i = p0 + p is used from 0,
// at the cell midpoint u = (x + 1/2) / 2^32 (never 0 or 1).
__global__ void k_normals_sobol_f64(const uint32_t* __restrict__ V, const uint32_t* __restrict__ shift,
                                    long p0, long np, int nstreams, int N, double* __restrict__ out) {
    const long dims = long(nstreams) * N;
    const long total = np * dims;
    for (long t = blockIdx.x * long(blockDim.x) + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x) {
        const long p = t / dims;
        const int d = int(t - p * dims);
        const uint64_t i = uint64_t(p0 + p + (shift ? 0 : 1));
        const uint32_t* v = V + size_t(d) * 32;
        uint32_t x = shift ? shift[d] : 0u;
        for (int k = 0; k < 32; ++k)
            if (i >> k & 1) x ^= v[k];
        const double u = (double(x) + (shift ? 0.5 : 0.0)) * 2.3283064365386963e-10;
        const int s = d % nstreams, j = d / nstreams;
        out[(p * nstreams + s) * N + j] = norm_inv_f64(u);
    }
}

hipError_t launch_normals_sobol_f64(const uint32_t* V, const uint32_t* shift, long p0, long np,
                                    int nstreams, int N, double* out, hipStream_t st) {
    const long total = np * long(nstreams) * N;
    hipLaunchKernelGGL(k_normals_sobol_f64, dim3(grid_for(total, 256)), dim3(256), 0, st, V, shift, p0,
                       np, nstreams, N, out);
    return hipGetLastError();
}

__global__ void k_norm_inv_probe(const double* __restrict__ u, long n, double* __restrict__ z) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i < n) z[i] = norm_inv_f64(u[i]);
}

hipError_t launch_norm_inv_probe(const double* u, long n, double* z, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_inv_probe, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, u, n, z);
    return hipGetLastError();
}

size_t volterra_f64_lds_bytes(int N) {
    const int nb = (N + 3) / 4;
    const int tpp = nb < 256 ? nb : 256;
    const int ppb = 256 / tpp;
    return size_t(N + ppb * 4 * (nb + 1)) * sizeof(double);
}

hipError_t launch_volterra_f64(const double* Z, long zstride, long np, int N, int nunits,
                               const Unit* units, const double* gamma, double* Wt,
                               hipStream_t st) {
    const int nb = (N + 3) / 4;
    const int tpp = nb < 256 ? nb : 256;
    const int ppb = 256 / tpp;
    const size_t lds = volterra_f64_lds_bytes(N);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_volterra_f64), lds);
        if (e != hipSuccess) return e;
    }
    dim3 grid(unsigned((np + ppb - 1) / ppb), unsigned(nunits));
    hipLaunchKernelGGL(k_volterra_f64, grid, dim3(256), lds, st, Z, zstride, np, N, units, gamma, Wt);
    return hipGetLastError();
}

hipError_t launch_scan_f64(int method, const double* Z, long zstride, const double* Wt, long np,
                           int N, const Group* groups, int ngroups, const double* comp,
                           double* out, hipStream_t st, double sign) {
    const long waves = np * ngroups;
    dim3 grid(unsigned((waves + 3) / 4));
    if (method == METHOD_RT)
        hipLaunchKernelGGL(k_scan_rt_f64, grid, dim3(256), 0, st, Z, zstride, Wt, np, N, groups,
                           ngroups, comp, out, sign);
    else
        hipLaunchKernelGGL(k_scan_mc_f64, grid, dim3(256), 0, st, Z, zstride, Wt, np, N, groups,
                           ngroups, comp, out);
    return hipGetLastError();
}

hipError_t launch_payoff_f64(const double* ints, long np, long p0, const Option* opts,
                             int out_size, int method, long blk0, long nblk_total,
                             double* partial, double* payoff, hipStream_t st) {
    dim3 grid(unsigned((np + 255) / 256), unsigned(out_size));
    hipLaunchKernelGGL(k_payoff_f64, grid, dim3(256), 0, st, ints, np, p0, opts, out_size, method,
                       blk0, nblk_total, partial, payoff);
    return hipGetLastError();
}

hipError_t launch_reduce_f64(const double* partial, long nblk, long M, int out_size,
                             double* price, double* stat, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_f64, dim3(out_size), dim3(256), 0, st, partial, nblk, 2, M, price,
                       stat);
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* partial, const long* row_off, const long* row_n,
                              const long* row_M, int nrows, double* price, double* stat,
                              hipStream_t st) {
    if (nrows < 1) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_rows, dim3(nrows), dim3(256), 0, st, partial, row_off, row_n, row_M,
                       price, stat);
    return hipGetLastError();
}

hipError_t launch_reduce_rows_iv(const double* partial, const long* row_off, const long* row_n,
                                 const long* row_M, const Option* opts, int nrows, double* price,
                                 double* stat, double* iv, int* counter, hipStream_t st) {
    if (nrows < 1) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_rows_iv, dim3(nrows), dim3(256), 0, st, partial, row_off, row_n, row_M,
                       opts, price, stat, iv, counter);
    return hipGetLastError();
}

hipError_t launch_iv_f64(const double* price, const Option* opts, int n, double* iv,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_iv_f64, dim3((n + 63) / 64), dim3(64), 0, st, price, opts, n, iv);
    return hipGetLastError();
}

hipError_t launch_iv_ref_f64(const double* price, const double* K, const double* T, int n, double* iv,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_iv_ref_f64, dim3((n + 63) / 64), dim3(64), 0, st, price, K, T, n, iv);
    return hipGetLastError();
}

hipError_t launch_philox_probe(const void* ctr, PhiloxKey key, long n, void* out, hipStream_t st) {
    hipLaunchKernelGGL(k_philox_probe, dim3(grid_for(n, 256)), dim3(256), 0, st,
                       static_cast<const uint4*>(ctr), key, n, static_cast<uint4*>(out));
    return hipGetLastError();
}

}  // namespace rb
