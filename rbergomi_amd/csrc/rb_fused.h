// rb_fused.h -- work tables of the fused mixed-precision kernel (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rb_device.h"

namespace rb {

constexpr int FUSED_MAX_UPW = 8;   // output row blocks (32 rows) per wave
constexpr int FUSED_GREG = 8;      // most groups integrated per pass over the accumulators
constexpr int FUSED_MAX_REG_OPTS = 256;  // options whose sums live in registers

struct FusedUnit {
    float b;         // W1perp weight of the first-cell term
    float sc;        // sqrt(2H) N^-H
    int frag_off;    // offset of this unit's A-fragment table, in 16-byte entries
    int p_off;       // offset of the unit's row tables, in floats.  Long-path kernel: one table
                     // P_j = (j - off)^(2H), P_j at j.  GEMM kernel: per group of the unit (in
                     // order, Np floats each) the row terms al P_j + be, at j - 1
    int g_begin, g_end;    // groups of this unit (contiguous)
    int pg_begin, pg_end;  // payoff groups (group, rho) of this unit (contiguous)
    int o_begin, o_end;    // options of this unit (contiguous)
};

// exponent of sqrt(v_j): log2 sqrt(v_j) = c1 Wtilde_{j-1} - (al P_j + be)
struct FusedGroup {
    float c1;        // eta T^H log2(e) / 2
    float al;        // eta^2 dt^(2H) log2(e) / 4
    float be;        // -ln(xi) log2(e) / 2
    int pg_begin, pg_end;  // payoff groups of this group (contiguous)
    int o_begin, o_end;    // options of this group (contiguous)
    double xi;
    double sxi;
    double dt;
    double sdt;
};

struct FusedPG {     // payoff group: one (group, rho)
    int group;       // global fused group index
    int pad;
    double rho;
};

struct FusedOption {
    int pg;          // global payoff-group index
    int out;         // output index I = h*op_size + t
    double K;
    double lnK;
};

// per (path, payoff group) values shared by all strikes (plain Monte Carlo: lnS = ln S_T and
// sig holds S_T itself)
struct PathPG {
    double lnS;      // ln S' = -rho^2 Iv / 2 + rho Is
    double sig;      // sqrt((1 - rho^2) Iv)
    float S;
    float invs;
    float lnSf;
    float sigf;
};

// One request of a batch: its tables, sizes and the split of its output row blocks over the
// four waves of a workgroup.
struct FusedReq {
    int N, Np, NK, PB, TP;       // steps, padded steps (x32), 16-step k-blocks, path blocks, paths per tile
    int log2TP;                  // log2(TP)
    int nunits, pgmax, out_size, gchunk, acc_in_lds;
    int log2slices;              // lanes per option in the payoff stage (register mode)
    int tiles_per_item;
    int mc;                      // 1 = plain Monte Carlo (wpart holds 3 sums per entry), 0 = RT
    long M, ntiles, nitems;
    PhiloxKey key;
    const FusedUnit* units;
    const FusedGroup* groups;
    const FusedPG* pgs;
    const FusedOption* opts;
    const uint4* frag;           // A fragments: per unit [Np/16][hi, lo][64 lanes] x 16 bytes
    const float* ptab;           // row terms al P_j + be, Np floats per group (FusedUnit::p_off)
    double* partial;             // [out_size][nitems][2]
    int wave_pb[4];
    int wave_nu[4];              // row blocks of each wave, listed in decreasing order
    int wave_rb[4][FUSED_MAX_UPW];
};

// A launch: the requests, the work items (one item = tiles_per_item tiles of one request, handed
// out through an atomic counter) and the LDS layout sized for the largest request.
struct FusedBatchArgs {
    const FusedReq* reqs;
    const int* item_req;         // request of each work item
    const int* item_local;       // item index within its request
    long total_items;
    int* counter;                // next work item (zeroed before the launch)
    int bf_u4;                   // LDS: B fragments, in 16-byte units
    int wpart_d;                 // LDS: cross-wave partials, in doubles
    int pgd_n;                   // LDS: per-(path, payoff group) records of one group chunk
    int acc_d;                   // LDS: option sums (largest request in LDS mode), in doubles
};

// Mills ratio M(z) = (1 - Phi(z)) / phi(z) for z >= 0, as M = t P(t), t = 1/(1 + z/3), with P a
// degree-10 polynomial fitted here (weighted least squares at Chebyshev nodes in t) to
// sqrt(pi/2) erfcx(z/sqrt 2); relative error <= 3.3e-7 in FP32 over z >= 0 (tests compare it
// with scipy.special.erfcx).  M(z) -> 1/z as z -> infinity.  t is the hardware reciprocal
// refined by one Newton step (within an ulp of 1/(1 + z/3); four instructions in place of the
// ten of a correctly rounded division); the denominator is capped at 1e30 so that z = inf gives
// M ~ 3e-31 rather than NaN (every caller multiplies M by exp(-z^2/2), which is 0 there).
__device__ __forceinline__ float mills_f32(float z) {
    const float c[11] = {3.333333433e-01f, 3.333306611e-01f, 2.964155674e-01f, 2.201747149e-01f, 1.412241310e-01f, -6.350896508e-02f, 2.205852419e-01f, -5.401814580e-01f, 4.626366198e-01f, -1.762671471e-01f, 2.557139099e-02f};
    const float d = fminf(fmaf(z, 0.333333333f, 1.0f), 1e30f);
    const float r = __builtin_amdgcn_rcpf(d);
    const float t = fmaf(r, fmaf(-d, r, 1.0f), r);
    float acc = c[10];
#pragma unroll
    for (int i = 9; i >= 0; --i) acc = fmaf(acc, t, c[i]);
    return acc * t;
}

// conditioned Black-Scholes payoff of one path for one strike (tau = 1, S0 = S').
// With E = S' phi(d1) = K phi(d2):  S' Phi(d1) = d1 >= 0 ? S' - E M(d1) : E M(-d1), and the
// same for K Phi(d2), so one exponential serves both normal CDFs; the three sign cases are
// selects (no divergent branch in the strike loop).  Deep out of the money (d1 < -8,
// Phi(d1) < 7e-16) the normal CDFs are evaluated in FP64, in one of two forms of the same value:
// Phi(e) = exp(-e^2/2) erfcx(-e/sqrt 2) / 2 with S' exp(-e1^2/2) = K exp(-e2^2/2), one
// exponential (the default: fewer instructions), or S' Phi(e1) - K Phi(e2) through erfc
// (LOWREG: fewer live registers, for the long-path kernel, which holds its Gaussians in
// registers; the erfcx form raised its register spills from 51 to 121 at N = 4096).
// Multiply-adds are written out (no implicit contraction), so every kernel instantiation that
// inlines this computes the same bits.
template <bool LOWREG = false>
__device__ __forceinline__ double payoff_mixed(const PathPG& q, float lnKf, float Kf, double lnK, double K) {
#pragma clang fp contract(off)
    const float S = q.S, sigf = q.sigf;
    const float d1 = fmaf(q.lnSf - lnKf, q.invs, 0.5f * sigf);
    const float d2 = d1 - sigf;
    const float E = __expf(fmaf(-0.5f * d1, d1, q.lnSf)) * 0.398942280f;
    const float a1 = mills_f32(fabsf(d1)), a2 = mills_f32(fabsf(d2));
    const float x_itm = fmaf(E, a2 - a1, S - Kf);   // d2 >= 0
    const float x_mid = fmaf(-E, a1 + a2, S);       // d1 >= 0 > d2
    const float x_otm = E * (a1 - a2);              // d1 < 0
    double x = double(d2 >= 0.f ? x_itm : (d1 >= 0.f ? x_mid : x_otm));
    if (d1 < -8.0f) {
        const double e1 = (q.lnS - lnK) / q.sig + 0.5 * q.sig, e2 = e1 - q.sig;
        if constexpr (LOWREG)
            x = exp(q.lnS) * phi_f64(e1) - K * phi_f64(e2);
        else
            x = 0.5 * exp(q.lnS - 0.5 * e1 * e1) *
                (erfcx(-0.7071067811865476 * e1) - erfcx(-0.7071067811865476 * e2));
    }
    return x;
}

// plain Monte Carlo payoff (S_T - K)^+ in FP64 (S_T in the sig slot)
__device__ __forceinline__ double payoff_mc(const PathPG& q, double K) {
    const double x = q.sig - K;
    return x > 0.0 ? x : 0.0;
}

size_t fused_req_lds_bytes(const FusedReq& a);
size_t fused_lds_bytes(const FusedBatchArgs& b);
// mode: 0 = FP16 hi+lo split, 1 = BF16 hi+lo split, 2 = FP32 operands
// mc: plain Monte Carlo kernel (every request of the batch), else Romano-Touzi; greg groups per
// pass: 1 to 6, or 8
hipError_t launch_fused(int mode, int upw, bool mc, int greg, const FusedBatchArgs& b, int grid,
                        hipStream_t st);
hipError_t launch_fused_rt(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt6(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt2(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt5(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc6(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc8(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc2(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc5(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt4(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc4(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt1(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc1(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_rt3(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
hipError_t launch_fused_mc3(int mode, int upw, const FusedBatchArgs& b, int grid, hipStream_t st);
// groups per pass of the fused kernel for units of at most gmax groups (1 to 6, or 8): the
// fewest passes (each pass reads the W1 tile again and runs the payoff stage over the tile's
// paths: N = 512, M = 2^22, five maturities took 9.10 ms in one pass and 11.5 / 14.1 / 19.4 ms
// in two / three / five), then the fewest groups per pass (a dead group slot costs a branch per
// 16 rows and holds accumulators in registers)
inline int fused_greg(int gmax) {
    const int passes = (gmax + FUSED_GREG - 1) / FUSED_GREG;
    for (int g : {1, 2, 3, 4, 5, 6})
        if ((gmax + g - 1) / g == passes) return g;
    return FUSED_GREG;
}
hipError_t launch_zero_int(int* p, int n, hipStream_t st);

// ---------------------------------------------------------------- long paths (N > 1024)
// One request priced by the single-launch FFT kernel (rb_fused_fft.hip).
struct FftFusedArgs {
    int N, log2L, nunits, out_size;
    int mc;                      // 1 = plain Monte Carlo, 0 = Romano-Touzi
    int pgu;                     // most payoff groups of one unit (LDS records per path)
    int pairs_per_item;          // path pairs of one work item
    long M, npairs, nitems;
    PhiloxKey key;
    const FusedUnit* units;      // b, sc, p_off and the group / payoff-group / option ranges
    const FusedGroup* groups;
    const FusedPG* pgs;
    const FusedOption* opts;
    const float* ptab;           // per unit: P_j = (j - off)^(2H)
    const float2* ghat;          // per unit [L]: FFT of (a, Gamma_1, ..., Gamma_{N-1}, 0, ...),
                                 // bit-reversed order
    const float2* tw;            // [L/2]: exp(-2 pi i k / L)
    double* partial;             // [out_size][nitems][2]
    int* counter;                // next work item (zeroed before the launch)
};
size_t fused_fft_lds_bytes(const FftFusedArgs& a);
// quads per thread of the kernel instantiation that runs N: ceil(ceil(N/4) / 256) rounded up to
// 1, 2, 4 or 8; its threads visit steps 0 .. 1024 QPT - 1
inline int fused_fft_qpt(int N) {
    const int q = ((N + 3) / 4 + 255) / 256;
    return q <= 1 ? 1 : q <= 2 ? 2 : q <= 4 ? 4 : 8;
}
hipError_t launch_fused_fft(const FftFusedArgs& a, int grid, hipStream_t st);

}  // namespace rb
