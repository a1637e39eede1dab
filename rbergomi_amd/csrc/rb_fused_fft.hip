// rb_fused_fft.hip -- single-launch mixed-precision Romano-Touzi pricing by FFT convolution: for
// long paths (N > 1024, where the Toeplitz GEMM's W1 tile no longer fits LDS), and for any
// N <= 8192 at the FFT precision (the same kernel with one quad per thread below N = 1024).
//
// A persistent workgroup takes work items (pairs_per_item pairs of paths) from an atomic
// counter and runs the whole path of the reference for each pair without touching HBM:
//   1. Philox4x32-10 W1 of both paths (FP32 Box-Muller, stream 0), kept in registers: thread t
//      owns the 4-step quads q = t + 256 j of both paths (genGaussianMT,
//      rBergLib/rBergomiMT.cpp:169-172).
//   2. Per unit (H): z = W1_a + i W1_b zero-padded to L = 2^k >= 2N - 1 in LDS, radix-16 FFT
//      (decimation in frequency), times FFT(a, Gamma_1, ..., Gamma_{N-1}) (the first-cell weight
//      a folded in; bit-reversed order), inverse FFT (decimation in time): Gamma is real, so the
//      real and imaginary parts are the two paths' convolutions (updateWtilde's FFTW convolution,
//      :83-109).  The transforms' top stages are done while z is written and read back.
//   3. Wtilde = sc (Y + b W1perp) for the thread's steps, W1perp regenerated from the same
//      counters (stream 1), written back to LDS in step order; then per (H, eta, T) group
//      sqrt(v_i) = exp2(c1 Wtilde_{i-1} - al P_i - be) and the sums of v and sqrt(v) W1 over
//      the thread's steps in FP32, over the workgroup in FP64 (scaleWtilde/updateV/intVdt/
//      intRootVdW, :73-80, 163-167, 427-436).
//   4. Per (path, group, rho) ln S', sigma; per option the conditioned Black-Scholes payoff in
//      FP32 with the FP64 tail (payoff_mixed), summed in FP64 per option over the item's paths.
// Plain Monte Carlo (computeIV_tensor's estimator, MC = true): a third sum sqrt(v) Wperp with
// Wperp from stream 2, ln S_T = rho Is + sqrt(1 - rho^2) Ip - Iv / 2 and the payoff (S_T - K)^+
// in FP64 (updateZ/updateS, rBergLib/rBergomiMT.cpp:111-137).
// Outputs the same per-item option sums as the GEMM kernel, reduced by k_reduce_rows.
#include <hip/hip_runtime.h>

#include "rb_fft_lds.h"
#include "rb_fused.h"
#include "rb_kernels.h"

namespace rb {

constexpr int FFT_GR = 8;  // groups integrated per pass

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int FFT_NS = 3;  // sums per (path, group) held in the reduction buffer

size_t fused_fft_lds_bytes(const FftFusedArgs& a) {
    return fft_lds_slots(a.log2L) * sizeof(float2) + size_t(4) * FFT_GR * 2 * FFT_NS * sizeof(double) +
           size_t(2) * a.pgu * sizeof(PathPG) + size_t(a.out_size) * 2 * sizeof(double);
}

// QPT: quads per thread and path, ceil(ceil(N/4) / 256); LOG2L: the transform length,
// L = 2^LOG2L >= 2N - 1 and at least 2^10 (LOG2L = 11 + log2 QPT for 1024 < N <= 8192; 10 or 11
// for QPT = 1).
template <int QPT, bool MC, int LOG2L>
__global__ void __launch_bounds__(256, (QPT >= 8 || (MC && QPT >= 4)) ? 1 : 2) k_fused_fft(FftFusedArgs A) {
    extern __shared__ float2 lds2[];
    const int L = 1 << A.log2L, H2 = L >> 1;
    float2* z = lds2;                                                      // [L + L/32], slot fpad(i)
    double* red = reinterpret_cast<double*>(z + L + (L >> 5));             // [4][FFT_GR][2][FFT_NS]
    PathPG* pgd = reinterpret_cast<PathPG*>(red + 4 * FFT_GR * 2 * FFT_NS);  // [2][pgu]
    double* acc = reinterpret_cast<double*>(pgd + 2 * A.pgu);              // [out][2]
    __shared__ int s_item;
    __shared__ float s_w10[2], s_w20[2];  // W1, Wperp of both paths at step 0
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int N = A.N, nq = (N + 3) >> 2;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(A.counter, 1);
        __syncthreads();
        const long item = s_item;
        __syncthreads();
        if (item >= A.nitems) break;
        for (int o = tid; o < 2 * A.out_size; o += 256) acc[o] = 0.0;
        const long pr0 = item * A.pairs_per_item;
        const long pr1 = pr0 + A.pairs_per_item < A.npairs ? pr0 + A.pairs_per_item : A.npairs;
        for (long pr = pr0; pr < pr1; ++pr) {
            const long pth[2] = {2 * pr, 2 * pr + 1};
            // ---- 1. W1 of both paths for this thread's quads (zero past N and past M)
            float w1[2][QPT][4];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int j = 0; j < QPT; ++j) {
                    const int q = tid + 256 * j;
                    float zz[4] = {0.f, 0.f, 0.f, 0.f};
                    if (q < nq && pth[c] < A.M) normals4_f32(A.key, uint64_t(pth[c]), 0u, uint32_t(q), zz);
#pragma unroll
                    for (int t = 0; t < 4; ++t) w1[c][j][t] = 4 * q + t < N ? zz[t] : 0.f;
                }
            if (tid == 0) {
                s_w10[0] = w1[0][0][0];
                s_w10[1] = w1[1][0][0];
                if constexpr (MC) {
                    for (int c = 0; c < 2; ++c) {
                        float zz[4] = {0.f, 0.f, 0.f, 0.f};
                        if (pth[c] < A.M) normals4_f32(A.key, uint64_t(pth[c]), 2u, 0u, zz);
                        s_w20[c] = zz[0];
                    }
                }
            }
            for (int u = 0; u < A.nunits; ++u) {
                const FusedUnit U = A.units[u];
                // ---- 2. z in natural order through the forward transform's top stage (its upper
                //         half is the zero padding: z_i = x_i, z_{i + L/2} = x_i W_L^i), the
                //         remaining stages, product, inverse transform but its top stage
#pragma unroll
                for (int j = 0; j < QPT; ++j) {
                    const int q = tid + 256 * j;
                    if (q < nq)
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int i = 4 * q + t;  // 4 nq <= L/2
                            const float2 x = make_float2(w1[0][j][t], w1[1][j][t]);
                            z[fpad(i)] = x;
                            z[fpad(i + H2)] = cmulf(x, A.tw[i]);
                        }
                }
                for (int i = 4 * nq + tid; i < H2; i += 256) {
                    z[fpad(i)] = make_float2(0.f, 0.f);
                    z[fpad(i + H2)] = make_float2(0.f, 0.f);
                }
                __syncthreads();
                fft_convolve_r16<LOG2L>(z, A.tw, A.ghat + long(u) * L, FftWorkgroup{tid});
                // ---- 3. Wtilde of the thread's steps (both paths), then back to LDS in step order
                const float invL = 1.0f / float(L);
                float wt[2][QPT][4];
#pragma unroll
                for (int j = 0; j < QPT; ++j) {
                    const int q = tid + 256 * j;
                    float zp[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        if (q < nq && pth[c] < A.M) normals4_f32(A.key, uint64_t(pth[c]), 1u, uint32_t(q), zp[c]);
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int i = 4 * q + t;
                        float2 y = make_float2(0.f, 0.f);
                        if (i < N) {  // the inverse top stage: z_i + conj(W_L^i) z_{i + L/2}
                            const float2 lo = z[fpad(i)], hi = z[fpad(i + H2)], w = A.tw[i];
                            const float2 p = cmulf(make_float2(w.x, -w.y), hi);
                            y = make_float2(lo.x + p.x, lo.y + p.y);
                        }
                        wt[0][j][t] = U.sc * fmaf(y.x, invL, U.b * zp[0][t]);
                        wt[1][j][t] = U.sc * fmaf(y.y, invL, U.b * zp[1][t]);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < QPT; ++j) {
                    const int q = tid + 256 * j;
                    if (q < nq)
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (4 * q + t < N) z[fpad(4 * q + t)] = make_float2(wt[0][j][t], wt[1][j][t]);
                }
                __syncthreads();
                // ---- 4. integrals of the unit's groups, FFT_GR groups per pass
                const float* P = A.ptab + U.p_off;
                for (int g0 = U.g_begin; g0 < U.g_end; g0 += FFT_GR) {
                    const int gn = U.g_end - g0 < FFT_GR ? U.g_end - g0 : FFT_GR;
                    float c1[FFT_GR], al[FFT_GR], be[FFT_GR];
                    float fv[FFT_GR][2], fs[FFT_GR][2], fp[FFT_GR][2];
#pragma unroll
                    for (int gg = 0; gg < FFT_GR; ++gg) {
                        c1[gg] = al[gg] = be[gg] = 0.f;
                        fv[gg][0] = fv[gg][1] = fs[gg][0] = fs[gg][1] = fp[gg][0] = fp[gg][1] = 0.f;
                        if (gg < gn) {
                            c1[gg] = A.groups[g0 + gg].c1;
                            al[gg] = A.groups[g0 + gg].al;
                            be[gg] = A.groups[g0 + gg].be;
                        }
                    }
                    if constexpr (MC) {
                        // plain MC: steps outside, groups inside, since the Wperp Gaussians come
                        // per quad (held across a group loop they raised the two-quad kernel's
                        // spills from 0 to 27 VGPRs)
#pragma unroll
                        for (int j = 0; j < QPT; ++j) {
                            const int q = tid + 256 * j;
                            float wp[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
                            for (int c = 0; c < 2; ++c)
                                if (q < nq && pth[c] < A.M) normals4_f32(A.key, uint64_t(pth[c]), 2u, uint32_t(q), wp[c]);
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                const int i = 4 * q + t;  // v_i uses Wtilde_{i-1}; i = 0 is v_0 = xi
                                if (i >= 1 && i < N) {
                                    const float2 x = z[fpad(i - 1)];
                                    const float pj = P[i];
#pragma unroll
                                    for (int gg = 0; gg < FFT_GR; ++gg) {
                                        if (gg < gn) {
                                            const float e = fmaf(al[gg], pj, be[gg]);
                                            const float ra = __builtin_amdgcn_exp2f(fmaf(x.x, c1[gg], -e));
                                            const float rb = __builtin_amdgcn_exp2f(fmaf(x.y, c1[gg], -e));
                                            fv[gg][0] = fmaf(ra, ra, fv[gg][0]);
                                            fs[gg][0] = fmaf(ra, w1[0][j][t], fs[gg][0]);
                                            fv[gg][1] = fmaf(rb, rb, fv[gg][1]);
                                            fs[gg][1] = fmaf(rb, w1[1][j][t], fs[gg][1]);
                                            fp[gg][0] = fmaf(ra, wp[0][t], fp[gg][0]);
                                            fp[gg][1] = fmaf(rb, wp[1][t], fp[gg][1]);
                                        }
                                    }
                                }
                            }
                        }
                    } else {
                        // Romano-Touzi: groups outside (one uniform branch per group), steps inside
                        // with no test:
                        // P = +inf at i = 0 and at i >= N (al > 0) makes those terms exactly 0, and
                        // i - 1 is clamped to a written slot.  Each group's sums still run over the
                        // steps in the same order.
#pragma unroll
                        for (int gg = 0; gg < FFT_GR; ++gg) {
                            if (gg < gn) {
#pragma unroll
                                for (int j = 0; j < QPT; ++j) {
                                    const int q = tid + 256 * j;
#pragma unroll
                                    for (int t = 0; t < 4; ++t) {
                                        const int i = 4 * q + t;  // v_i uses Wtilde_{i-1}; i = 0 is v_0 = xi
                                        const float2 x = z[fpad(i > 0 ? i - 1 : 0)];
                                        const float e = fmaf(al[gg], P[i], be[gg]);
                                        const float ra = __builtin_amdgcn_exp2f(fmaf(x.x, c1[gg], -e));
                                        const float rb = __builtin_amdgcn_exp2f(fmaf(x.y, c1[gg], -e));
                                        fv[gg][0] = fmaf(ra, ra, fv[gg][0]);
                                        fs[gg][0] = fmaf(ra, w1[0][j][t], fs[gg][0]);
                                        fv[gg][1] = fmaf(rb, rb, fv[gg][1]);
                                        fs[gg][1] = fmaf(rb, w1[1][j][t], fs[gg][1]);
                                    }
                                }
                            }
                        }
                    }
                    // FP64 over the workgroup: wave sums, then the four waves in order
#pragma unroll
                    for (int gg = 0; gg < FFT_GR; ++gg) {
                        if (gg < gn) {
#pragma unroll
                            for (int c = 0; c < 2; ++c) {
                                const double sv = wave_sum64(double(fv[gg][c]));
                                const double ss = wave_sum64(double(fs[gg][c]));
                                double sp = 0.0;
                                if constexpr (MC) sp = wave_sum64(double(fp[gg][c]));
                                if (lane == 0) {
                                    red[((wave * FFT_GR + gg) * 2 + c) * FFT_NS + 0] = sv;
                                    red[((wave * FFT_GR + gg) * 2 + c) * FFT_NS + 1] = ss;
                                    red[((wave * FFT_GR + gg) * 2 + c) * FFT_NS + 2] = sp;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    if (tid < 2 * gn) {
#pragma clang fp contract(off)
                        const int gg = tid >> 1, c = tid & 1;
                        const FusedGroup Gr = A.groups[g0 + gg];
                        double a = Gr.xi, b = Gr.sxi * double(s_w10[c]);  // step 0: v_0 = xi
                        double cp = MC ? Gr.sxi * double(s_w20[c]) : 0.0;
                        for (int w = 0; w < 4; ++w) {
                            a += red[((w * FFT_GR + gg) * 2 + c) * FFT_NS + 0];
                            b += red[((w * FFT_GR + gg) * 2 + c) * FFT_NS + 1];
                            if constexpr (MC) cp += red[((w * FFT_GR + gg) * 2 + c) * FFT_NS + 2];
                        }
                        const double Iv = Gr.dt * a, Is = Gr.sdt * b;
                        if constexpr (MC) {
                            // ln S_T = rho Is + sqrt(1 - rho^2) Ip - Iv / 2 (log-Euler, left point)
                            const double Ip = Gr.sdt * cp;
                            for (int pg = Gr.pg_begin; pg < Gr.pg_end; ++pg) {
                                const double rho = A.pgs[pg].rho;
                                PathPG rec;
                                rec.lnS = rho * Is + sqrt(1.0 - rho * rho) * Ip - 0.5 * Iv;
                                rec.sig = exp(rec.lnS);
                                rec.S = float(rec.sig);
                                rec.ik = rec.A = rec.sg = 0.f;
                                pgd[c * A.pgu + (pg - U.pg_begin)] = rec;
                            }
                        }
                        for (int pg = Gr.pg_begin; !MC && pg < Gr.pg_end; ++pg) {
                            const double rho = A.pgs[pg].rho;
                            PathPG rec;
                            rec.sig = sqrt((1.0 - rho * rho) * Iv);
                            rec.lnS = -0.5 * rho * rho * Iv + rho * Is;
                            pathpg_f32(rec);
                            pgd[c * A.pgu + (pg - U.pg_begin)] = rec;
                        }
                    }
                    __syncthreads();
                }
                // ---- 5. payoffs of the unit's options for both paths (one thread per option)
                for (int o = U.o_begin + tid; o < U.o_end; o += 256) {
                    const FusedOption O = A.opts[o];
                    const OptF of = opt_f32(O.lnK, O.K);
                    double s1 = 0.0, s2 = 0.0;
                    for (int c = 0; c < 2; ++c) {
                        if (pth[c] >= A.M) continue;
                        const PathPG& rec = pgd[c * A.pgu + (O.pg - U.pg_begin)];
                        const double x = MC ? payoff_mc(rec, O.K) : payoff_mixed<true>(rec, of, O.lnK, O.K);
                        s1 += x;
                        s2 += x * x;
                    }
                    acc[2 * o + 0] += s1;
                    acc[2 * o + 1] += s2;
                }
                __syncthreads();
            }
        }
        for (int o = tid; o < A.out_size; o += 256) {
            const int I = A.opts[o].out;
            A.partial[(long(I) * A.nitems + item) * 2 + 0] = acc[2 * o + 0];
            A.partial[(long(I) * A.nitems + item) * 2 + 1] = acc[2 * o + 1];
        }
        __syncthreads();
    }
}

hipError_t launch_fused_fft(const FftFusedArgs& a, int grid, hipStream_t st) {
    const size_t lds = fused_fft_lds_bytes(a);
    const int qpt = fused_fft_qpt(a.N);
    // the kernels are compiled for the transform length of their quad count
    const int want = qpt == 1 ? (a.N <= 512 ? 10 : 11) : qpt == 2 ? 12 : qpt == 4 ? 13 : 14;
    if (a.N < 1 || a.log2L != want) return hipErrorInvalidValue;
    auto go = [&](auto kern) -> hipError_t {
        if (lds > 65536) {
            const hipError_t e = allow_lds(reinterpret_cast<const void*>(kern), lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, st, a);
        return hipGetLastError();
    };
    if (a.mc) {
        if (qpt == 1) return a.log2L == 10 ? go(&k_fused_fft<1, true, 10>) : go(&k_fused_fft<1, true, 11>);
        if (qpt <= 2) return go(&k_fused_fft<2, true, 12>);
        if (qpt <= 4) return go(&k_fused_fft<4, true, 13>);
        if (qpt <= 8) return go(&k_fused_fft<8, true, 14>);
    } else {
        if (qpt == 1) return a.log2L == 10 ? go(&k_fused_fft<1, false, 10>) : go(&k_fused_fft<1, false, 11>);
        if (qpt <= 2) return go(&k_fused_fft<2, false, 12>);
        if (qpt <= 4) return go(&k_fused_fft<4, false, 13>);
        if (qpt <= 8) return go(&k_fused_fft<8, false, 14>);
    }
    return hipErrorInvalidValue;
}

}  // namespace rb
