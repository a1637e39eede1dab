// rb_fused.h -- work tables of the fused mixed-precision kernel (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rb_device.h"

namespace rb {

// 16-byte units per 16-step chunk of a path block in the LDS W1 tile: the FP16 mode keeps one
// FP16 value per step (64 lane slots), the BF16 (hi, lo) and FP32 modes two halves (128)
constexpr int fused_chunk_u4(int mode) { return mode == 0 ? 64 : 128; }
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
// sig holds S_T itself, the FP32 fields are unused).  The FP32 fields give the conditioned
// Black-Scholes d1, d2 scaled by s = sqrt(log2(e) / 2), so that K phi(d2) is one exp2:
// s d2 = A - ln K ik, s d1 = s d2 + sg, K phi(d2) = exp2(log2(K / sqrt(2 pi)) - (s d2)^2).
struct PathPG {
    double lnS;      // ln S' = -rho^2 Iv / 2 + rho Is
    double sig;      // sqrt((1 - rho^2) Iv)
    float S;         // S'
    float ik;        // s / sig
    float A;         // s (ln S' / sig - sig / 2)
    float sg;        // s sig
};

constexpr double PAYOFF_S = 0.8493218002880191;   // sqrt(log2(e) / 2)

// the FP32 fields of a Romano-Touzi record from its FP64 ln S' and sigma
__device__ __forceinline__ void pathpg_f32(PathPG& q) {
    q.S = float(exp(q.lnS));
    q.ik = float(PAYOFF_S / q.sig);
    q.A = float(PAYOFF_S * (q.lnS / q.sig - 0.5 * q.sig));
    q.sg = float(PAYOFF_S * q.sig);
}

// per-strike FP32 constants of the mixed payoff
struct OptF {
    float lnK;       // ln K
    float K;
    float LK;        // log2(K / sqrt(2 pi))
};

__device__ __forceinline__ OptF opt_f32(double lnK, double K) {
    OptF o;
    o.lnK = float(lnK);
    o.K = float(K);
    o.LK = float(lnK * 1.4426950408889634 - 1.3257480647361595);
    return o;
}

// One request of a batch: its tables, sizes and the split of its output row blocks over the
// four waves of a workgroup.
struct FusedReq {
    int N, Np, NK, PB, TP;       // steps, padded steps (x32), 16-step k-blocks, path blocks, paths per tile
    int log2TP;                  // log2(TP)
    int nunits, pgmax, out_size, gchunk, acc_in_lds;
    int log2slices;              // lanes per option in the payoff stage (register mode)
    int pay2;                    // with one lane per option: options 0 .. pay2 - 1 take two (whole waves)
    int tiles_per_item;          // tiles of items 0 .. items_big - 1
    int mc;                      // 1 = plain Monte Carlo (wpart holds 3 sums per entry), 0 = RT
    long M, ntiles, nitems;
    long items_big;              // items of tiles_per_item tiles; the rest have tiles_small
    int tiles_small;
    int chunk_u4;                // fused_chunk_u4 of the launch's operand mode
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
// sqrt(pi/2) erfcx(z/sqrt 2); relative error <= 3.5e-7 in FP32 over z >= 0 with t rounded
// correctly, <= 4.9e-7 with t from the hardware reciprocal (within an ulp; tests compare it
// with scipy.special.erfcx).  M(z) -> 1/z as z -> infinity; z = inf gives t = 0 and M = 0.
// mills_den takes the denominator 1 + z/3 itself.
__device__ __forceinline__ float mills_den(float d) {
    const float c[11] = {3.333333433e-01f, 3.333306611e-01f, 2.964155674e-01f, 2.201747149e-01f, 1.412241310e-01f, -6.350896508e-02f, 2.205852419e-01f, -5.401814580e-01f, 4.626366198e-01f, -1.762671471e-01f, 2.557139099e-02f};
    const float t = __builtin_amdgcn_rcpf(d);
    float acc = c[10];
#pragma unroll
    for (int i = 9; i >= 0; --i) acc = fmaf(acc, t, c[i]);
    return acc * t;
}

__device__ __forceinline__ float mills_f32(float z) { return mills_den(fmaf(z, 0.333333333f, 1.0f)); }

// conditioned Black-Scholes payoff of one path for one strike (tau = 1, S0 = S').
// With E = S' phi(d1) = K phi(d2):  S' Phi(d1) = d1 >= 0 ? S' - E M(d1) : E M(-d1), and the
// same for K Phi(d2), so one exponential serves both normal CDFs; the three sign cases are
// selects (no divergent branch in the strike loop).  In FP32 from the record's scaled fields
// (u = s d, s = sqrt(log2(e) / 2)): u2 = A - ln K ik, u1 = u2 + sg, E = exp2(LK - u2^2), and
// the Mills-ratio denominators 1 + |u| / (3 s).  Deep out of the money (d1 < -8,
// Phi(d1) < 7e-16) the normal CDFs are evaluated in FP64, in one of two forms of the same value:
// Phi(e) = exp(-e^2/2) erfcx(-e/sqrt 2) / 2 with S' exp(-e1^2/2) = K exp(-e2^2/2), one
// exponential (the default: fewer instructions), or S' Phi(e1) - K Phi(e2) through erfc
// (LOWREG: fewer live registers, for the long-path kernel, which holds its Gaussians in
// registers; the erfcx form raised its register spills from 51 to 121 at N = 4096).
// Multiply-adds are written out (no implicit contraction), so every kernel instantiation that
// inlines this computes the same bits.
// the FP32 part; deep: d1 < -8, where the caller takes payoff_deep instead
__device__ __forceinline__ double payoff_f32(const PathPG& q, const OptF& o, bool& deep) {
#pragma clang fp contract(off)
    const float u2 = fmaf(-o.lnK, q.ik, q.A);
    const float u1 = u2 + q.sg;
    const float E = __builtin_amdgcn_exp2f(fmaf(-u2, u2, o.LK));
    const float k3 = float(1.0 / (3.0 * PAYOFF_S));
    const float a1 = mills_den(fmaf(fabsf(u1), k3, 1.0f)), a2 = mills_den(fmaf(fabsf(u2), k3, 1.0f));
    const float x_itm = fmaf(E, a2 - a1, q.S - o.K);   // d2 >= 0
    const float x_mid = fmaf(-E, a1 + a2, q.S);        // d1 >= 0 > d2
    const float x_otm = -E * (a2 - a1);                // d1 < 0: E (a1 - a2), same bits
    deep = u1 < float(-8.0 * PAYOFF_S);
    return double(u2 >= 0.f ? x_itm : (u1 >= 0.f ? x_mid : x_otm));
}

// the FP64 deep out-of-the-money form
template <bool LOWREG = false>
__device__ __forceinline__ double payoff_deep(const PathPG& q, double lnK, double K) {
#pragma clang fp contract(off)
    const double e1 = (q.lnS - lnK) / q.sig + 0.5 * q.sig, e2 = e1 - q.sig;
    if constexpr (LOWREG)
        return exp(q.lnS) * phi_f64(e1) - K * phi_f64(e2);
    else
        return 0.5 * exp(q.lnS - 0.5 * e1 * e1) *
               (erfcx(-0.7071067811865476 * e1) - erfcx(-0.7071067811865476 * e2));
}

template <bool LOWREG = false>
__device__ __forceinline__ double payoff_mixed(const PathPG& q, const OptF& o, double lnK, double K) {
    bool deep;
    const double x = payoff_f32(q, o, deep);
    return deep ? payoff_deep<LOWREG>(q, lnK, K) : x;
}

// plain Monte Carlo payoff (S_T - K)^+ in FP64 (S_T in the sig slot)
__device__ __forceinline__ double payoff_mc(const PathPG& q, double K) {
    const double x = q.sig - K;
    return x > 0.0 ? x : 0.0;
}

size_t fused_req_lds_bytes(const FusedReq& a);
size_t fused_lds_bytes(const FusedBatchArgs& b);
// mode: 0 = FP16 (Gamma split hi + lo, W1 rounded to FP16), 1 = BF16 hi + lo split of both, 2 = FP32
// operands
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
