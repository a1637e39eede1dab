// rb_fused.hip -- fused mixed-precision pricing kernel for N <= 1024 (gfx950).
//
// One workgroup (4 waves) owns a tile of TP = 32*PB paths at a time and runs the whole path
// of the reference for it without touching HBM in between:
//   1. Philox4x32-10 Gaussians W1 (FP32 Box-Muller) -> LDS, split hi + lo (FP16 or BF16) and
//      stored directly in the B-operand fragment layout of v_mfma_f32_32x32x16_{f16,bf16}
//      (replaces genGaussianMT, rBergLib/rBergomiMT.cpp:169-172)
//   2. Y' = L W1 on the matrix cores, L the lower-triangular Toeplitz matrix of Gamma with the
//      first-cell weight on its diagonal, as three products hi*hi + hi*lo + lo*hi accumulated
//      in FP32 (replaces the FFTW convolution of updateWtilde, :83-109).  The A operand of a
//      32x16 block depends only on its diagonal offset d, so A fragments come from a small
//      per-H table (no N x N matrix).  A wave walks d in phases over which the set of its
//      active row blocks is fixed, so the MFMAs of different row blocks interleave.
//   3. Epilogue in registers: Wtilde = sc (Y' + b W1perp) with W1perp regenerated from the
//      same counters, then for every (H, eta, T) group sqrt(v_{i+1}) = exp2(c1 Wtilde_i -
//      al P_{i+1} - be) and the sums of v and sqrt(v) W1 (scaleWtilde/updateV/intVdt/
//      intRootVdW, :73-80, 163-167, 427-436): FP32 terms summed per 16-row block, FP64 beyond.
//   4. Per (path, group, rho): ln S', sigma, 1/sigma, S' once; per strike the conditioned
//      Black-Scholes payoff S' Phi(d1) - K Phi(d2) (:383-391) in FP32, with the FP64 normal
//      CDF for deep out-of-the-money lanes (d1 < -8), for the options of up to 8 groups at a
//      time.  The paths of an option are split over S = 2^k lanes (S ~ 256 / options) and summed
//      in FP64 with a fixed butterfly into the register of the option's owner lane.
#include <hip/hip_runtime.h>

#include "rb_fused.h"
#include "rb_kernels.h"
#include "rb_mfma.h"

namespace rb {

template <int MODE>
struct SplitT;

template <>
struct SplitT<0> {
    typedef _Float16 T;
    typedef half8_t V;
    typedef _Float16 V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ T cvt(float x) { return (T)x; }
    static __device__ __forceinline__ floatx16_t mma(V a, V b, floatx16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct SplitT<1> {
    typedef __bf16 T;
    typedef bf16x8_t V;
    typedef __bf16 V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ T cvt(float x) { return (T)x; }
    static __device__ __forceinline__ floatx16_t mma(V a, V b, floatx16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ double half_sum_f64(double v) {
    // sum over the 32 lanes of one half-wave; every lane of the half receives the total
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int MODE, int UPW>
__global__ void __launch_bounds__(256, UPW >= 8 ? 1 : 2) k_fused(FusedBatchArgs B) {
    typedef SplitT<MODE> S;
    typedef typename S::T T;
    typedef typename S::V V;
    typedef typename S::V4 V4;
    extern __shared__ uint4 lds4[];
    __shared__ int s_item;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, col = lane & 31;
    uint4* Bf = lds4;                                                   // [PB][NK][2][64]
    double* wpart = reinterpret_cast<double*>(lds4 + B.bf_u4);          // [4][gchunk][32][2]
    PathPG* pgd = reinterpret_cast<PathPG*>(wpart + B.wpart_d);         // [TP][pgmax]
    double* accopt = reinterpret_cast<double*>(pgd + B.pgd_n);          // [out][2] in LDS mode

    for (;;) {
        // ---- next work item (dynamic schedule; the item's sums do not depend on who runs it)
        if (tid == 0) s_item = atomicAdd(B.counter, 1);
        __syncthreads();
        const long gi = s_item;
        __syncthreads();
        if (gi >= B.total_items) break;
        const FusedReq A = B.reqs[B.item_req[gi]];   // copied once per item (no aliasing reloads)
        const long item = B.item_local[gi];
        const int NK = A.NK;
        const int pbw = A.wave_pb[wave];
        const int nuw = A.wave_nu[wave];
        double a1 = 0.0, a2 = 0.0;  // sums of option tid (register accumulation)
        if (A.acc_in_lds)
            for (int o = tid; o < 2 * A.out_size; o += 256) accopt[o] = 0.0;
        const long t_begin = item * A.tiles_per_item;
        const long t_end = t_begin + A.tiles_per_item < A.ntiles ? t_begin + A.tiles_per_item : A.ntiles;
        for (long tile = t_begin; tile < t_end; ++tile) {
            const long p0 = tile * A.TP;
            // ---- 1. Gaussians W1 -> B fragments (hi, lo); zero past N and past M
            const int nq = A.Np >> 2;
            {
                const int pl = tid & (A.TP - 1);   // TP is 32, 64 or 128
                const int pb = pl >> 5;
                const long p = p0 + pl;
                const bool live = p < A.M;
                for (int q = tid >> A.log2TP; q < nq; q += 256 >> A.log2TP) {
                    float z[4] = {0.f, 0.f, 0.f, 0.f};
                    if (live && 4 * q < A.N) {
                        normals4_f32(A.key, uint64_t(p), 0u, uint32_t(q), z);
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (4 * q + j >= A.N) z[j] = 0.f;
                    }
                    V4 hi, lo;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        hi[j] = S::cvt(z[j]);
                        lo[j] = S::cvt(z[j] - float(hi[j]));
                    }
                    const int ks = q >> 2, hh = (q & 3) >> 1, j0 = (q & 1);
                    const int ln = 32 * hh + (pl & 31);
                    uint4* base = Bf + ((pb * NK + ks) * 2) * 64;
                    reinterpret_cast<uint2*>(base + ln)[j0] = __builtin_bit_cast(uint2, hi);
                    reinterpret_cast<uint2*>(base + 64 + ln)[j0] = __builtin_bit_cast(uint2, lo);
                }
            }
            __syncthreads();

            for (int u = 0; u < A.nunits; ++u) {
                const FusedUnit U = A.units[u];
                // ---- 2. Toeplitz GEMM on the matrix cores
                floatx16_t acc[UPW];
#pragma unroll
                for (int s = 0; s < UPW; ++s)
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[s][r] = 0.f;
                if (pbw >= 0) {
                    const uint4* F = A.frag + U.frag_off;
                    // phase k: row blocks 0..k-1 of the wave are active (blocks sorted descending)
#pragma unroll
                    for (int k = 1; k <= UPW; ++k) {
                        if (k > nuw) break;
                        const int dhi = 32 * A.wave_rb[wave][k - 1];
                        const int dlo = k < nuw ? 32 * A.wave_rb[wave][k] + 16 : -16;
                        for (int d = dhi; d >= dlo; d -= 16) {
                            const uint4* Fd = F + ((d + 16) >> 4) * 128;
                            const V ahi = __builtin_bit_cast(V, Fd[lane]);
                            const V alo = __builtin_bit_cast(V, Fd[64 + lane]);
                            V bhi[UPW], blo[UPW];
#pragma unroll
                            for (int s = 0; s < k; ++s) {
                                const int ks = (32 * A.wave_rb[wave][s] - d) >> 4;
                                const uint4* Bp = Bf + ((pbw * NK + ks) * 2) * 64;
                                bhi[s] = __builtin_bit_cast(V, Bp[lane]);
                                blo[s] = __builtin_bit_cast(V, Bp[64 + lane]);
                            }
#pragma unroll
                            for (int s = 0; s < k; ++s) acc[s] = S::mma(ahi, bhi[s], acc[s]);
#pragma unroll
                            for (int s = 0; s < k; ++s) acc[s] = S::mma(ahi, blo[s], acc[s]);
#pragma unroll
                            for (int s = 0; s < k; ++s) acc[s] = S::mma(alo, bhi[s], acc[s]);
                        }
                    }
                    // ---- 3a. Wtilde = sc (Y' + b W1perp), W1perp from the same counters (stream 1)
                    const long p = p0 + 32 * pbw + col;
#pragma unroll
                    for (int s = 0; s < UPW; ++s) {
                        if (s >= nuw) break;
                        const int rb = A.wave_rb[wave][s];
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            float z[4] = {0.f, 0.f, 0.f, 0.f};
                            if (p < A.M) normals4_f32(A.key, uint64_t(p), 1u, uint32_t(8 * rb + 2 * c + h), z);
#pragma unroll
                            for (int t = 0; t < 4; ++t)
                                acc[s][4 * c + t] = U.sc * (acc[s][4 * c + t] + U.b * z[t]);
                        }
                    }
                }
                // ---- 3b. integrals of every group of the unit, FUSED_GREG groups per pass
                for (int g0 = U.g_begin; g0 < U.g_end; g0 += FUSED_GREG) {
                    const int gn = U.g_end - g0 < FUSED_GREG ? U.g_end - g0 : FUSED_GREG;
                    float c1[FUSED_GREG], al[FUSED_GREG], be[FUSED_GREG];
                    double sv[FUSED_GREG], ss[FUSED_GREG];
#pragma unroll
                    for (int gg = 0; gg < FUSED_GREG; ++gg) {
                        sv[gg] = ss[gg] = 0.0;
                        c1[gg] = al[gg] = be[gg] = 0.f;
                        if (gg < gn) {
                            c1[gg] = A.groups[g0 + gg].c1;
                            al[gg] = A.groups[g0 + gg].al;
                            be[gg] = A.groups[g0 + gg].be;
                        }
                    }
                    if (pbw >= 0) {
                        const float* P = A.ptab + U.p_off;
#pragma unroll
                        for (int s = 0; s < UPW; ++s) {
                            if (s >= nuw) break;
                            const int rb = A.wave_rb[wave][s];
                            float fv[FUSED_GREG], fs[FUSED_GREG];
#pragma unroll
                            for (int gg = 0; gg < FUSED_GREG; ++gg) fv[gg] = fs[gg] = 0.f;
#pragma unroll
                            for (int r = 0; r < 16; ++r) {
                                const int j = 32 * rb + mfma32_row(r, lane) + 1;  // v_j uses Wtilde_{j-1}
                                if (j < A.N) {
                                    const int ks = j >> 4, kk = j & 15;
                                    const uint4* Bp = Bf + ((pbw * NK + ks) * 2) * 64 + 32 * (kk >> 3) + col;
                                    const float w1 = float(reinterpret_cast<const T*>(Bp)[kk & 7]) +
                                                     float(reinterpret_cast<const T*>(Bp + 64)[kk & 7]);
                                    const float pj = P[j];
                                    const float x = acc[s][r];
#pragma unroll
                                    for (int gg = 0; gg < FUSED_GREG; ++gg) {
                                        if (gg < gn) {
                                            const float rr = __builtin_amdgcn_exp2f(fmaf(x, c1[gg], -fmaf(al[gg], pj, be[gg])));
                                            fv[gg] = fmaf(rr, rr, fv[gg]);
                                            fs[gg] = fmaf(rr, w1, fs[gg]);
                                        }
                                    }
                                }
                            }
#pragma unroll
                            for (int gg = 0; gg < FUSED_GREG; ++gg) {
                                sv[gg] += double(fv[gg]);
                                ss[gg] += double(fs[gg]);
                            }
                        }
                    }
#pragma unroll
                    for (int gg = 0; gg < FUSED_GREG; ++gg) {
                        sv[gg] += __shfl_xor(sv[gg], 32, 64);
                        ss[gg] += __shfl_xor(ss[gg], 32, 64);
                    }
                    // cross-wave sums in chunks of gchunk groups (LDS budget)
                    const int pg_lo = A.groups[g0].pg_begin;
                    for (int c0 = 0; c0 < gn; c0 += A.gchunk) {
                        const int cn = gn - c0 < A.gchunk ? gn - c0 : A.gchunk;
                        if (h == 0) {
#pragma unroll
                            for (int gg = 0; gg < FUSED_GREG; ++gg)
                                if (gg >= c0 && gg < c0 + cn) {
                                    double* w = wpart + ((wave * A.gchunk + (gg - c0)) * 32 + col) * 2;
                                    w[0] = sv[gg];
                                    w[1] = ss[gg];
                                }
                        }
                        __syncthreads();
                        for (int e = tid; e < A.TP * cn; e += 256) {
                            const int pl = e & (A.TP - 1), gg = e >> A.log2TP;
                            const int pb = pl >> 5, c = pl & 31;
                            const FusedGroup G = A.groups[g0 + c0 + gg];
                            const uint4* B0 = Bf + ((pb * NK) * 2) * 64 + c;  // k = 0
                            const double w10 = double(float(reinterpret_cast<const T*>(B0)[0]) +
                                                      float(reinterpret_cast<const T*>(B0 + 64)[0]));
                            double a = G.xi, b = G.sxi * w10;   // j = 0: v_0 = xi
                            for (int w = 0; w < 4; ++w)
                                if (A.wave_pb[w] == pb) {
                                    a += wpart[((w * A.gchunk + gg) * 32 + c) * 2 + 0];
                                    b += wpart[((w * A.gchunk + gg) * 32 + c) * 2 + 1];
                                }
                            const double Iv = G.dt * a, Is = G.sdt * b;
                            for (int pg = G.pg_begin; pg < G.pg_end; ++pg) {
                                const double rho = A.pgs[pg].rho;
                                PathPG q;
                                q.sig = sqrt((1.0 - rho * rho) * Iv);
                                q.lnS = -0.5 * rho * rho * Iv + rho * Is;
                                q.S = float(exp(q.lnS));
                                q.invs = float(1.0 / q.sig);
                                q.lnSf = float(q.lnS);
                                q.sigf = float(q.sig);
                                pgd[pl * A.pgmax + (pg - pg_lo)] = q;
                            }
                        }
                        __syncthreads();
                    }
                    // ---- 4. payoffs of the options of groups g0 .. g0+gn-1.  Register mode: the
                    //         S = 2^log2slices lanes t = o*S .. o*S+S-1 share option o, each
                    //         summing every S-th path; a fixed butterfly leaves the total in lane
                    //         o*S, which owns option o's FP64 sums.  LDS mode (> 256 options):
                    //         one thread per option, sums in LDS.
                    const int o_lo = A.groups[g0].o_begin, o_hi = A.groups[g0 + gn - 1].o_end;
                    const int np = A.M - p0 < A.TP ? int(A.M - p0) : A.TP;
                    if (A.acc_in_lds) {
                        for (int oi = o_lo + tid; oi < o_hi; oi += 256) {
                            const FusedOption O = A.opts[oi];
                            const int pgl = O.pg - pg_lo;
                            const float lnKf = float(O.lnK), Kf = float(O.K);
                            double s1 = 0.0, s2 = 0.0;
                            for (int pl = 0; pl < np; ++pl) {
                                const double x = payoff_mixed(pgd[pl * A.pgmax + pgl], lnKf, Kf, O.lnK, O.K);
                                s1 += x;
                                s2 += x * x;
                            }
                            accopt[2 * oi + 0] += s1;
                            accopt[2 * oi + 1] += s2;
                        }
                    } else {
                        const int S = 1 << A.log2slices;
                        const int oi = tid >> A.log2slices, sl = tid & (S - 1);
                        if (oi >= o_lo && oi < o_hi) {
                            const FusedOption O = A.opts[oi];
                            const int pgl = O.pg - pg_lo;
                            const float lnKf = float(O.lnK), Kf = float(O.K);
                            double s1 = 0.0, s2 = 0.0;
                            for (int pl = sl; pl < np; pl += S) {
                                const double x = payoff_mixed(pgd[pl * A.pgmax + pgl], lnKf, Kf, O.lnK, O.K);
                                s1 += x;
                                s2 += x * x;
                            }
                            for (int off = S >> 1; off >= 1; off >>= 1) {
                                s1 += __shfl_xor(s1, off, 64);
                                s2 += __shfl_xor(s2, off, 64);
                            }
                            if (sl == 0) {
                                a1 += s1;
                                a2 += s2;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
        }
        // ---- item partial sums (indexed by the output index I)
        if (A.acc_in_lds) {
            for (int o = tid; o < A.out_size; o += 256) {
                const int I = A.opts[o].out;
                A.partial[(long(I) * A.nitems + item) * 2 + 0] = accopt[2 * o + 0];
                A.partial[(long(I) * A.nitems + item) * 2 + 1] = accopt[2 * o + 1];
            }
        } else if ((tid & ((1 << A.log2slices) - 1)) == 0 && (tid >> A.log2slices) < A.out_size) {
            const int I = A.opts[tid >> A.log2slices].out;
            A.partial[(long(I) * A.nitems + item) * 2 + 0] = a1;
            A.partial[(long(I) * A.nitems + item) * 2 + 1] = a2;
        }
        __syncthreads();
    }
}

size_t fused_req_lds_bytes(const FusedReq& a) {
    return size_t(a.PB) * a.NK * 128 * 16 + size_t(4) * a.gchunk * 32 * 2 * 8 +
           size_t(a.TP) * a.pgmax * sizeof(PathPG) + (a.acc_in_lds ? size_t(a.out_size) * 2 * 8 : 0);
}

size_t fused_lds_bytes(const FusedBatchArgs& b) {
    return size_t(b.bf_u4) * 16 + size_t(b.wpart_d) * 8 + size_t(b.pgd_n) * sizeof(PathPG) +
           size_t(b.acc_d) * 8;
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_int(int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

template <int MODE, int UPW>
static hipError_t launch_t(const FusedBatchArgs& a, int grid, hipStream_t st) {
    const size_t lds = fused_lds_bytes(a);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_fused<MODE, UPW>), lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_fused<MODE, UPW>), dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_m(int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    switch (upw) {
        case 1: return launch_t<MODE, 1>(a, grid, st);
        case 2: return launch_t<MODE, 2>(a, grid, st);
        case 4: return launch_t<MODE, 4>(a, grid, st);
        case 8: return launch_t<MODE, 8>(a, grid, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fused(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    return mode == 0 ? launch_m<0>(upw, a, grid, st) : launch_m<1>(upw, a, grid, st);
}

}  // namespace rb
