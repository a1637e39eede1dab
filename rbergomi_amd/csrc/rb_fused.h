// rb_fused.h -- work tables of the fused mixed-precision kernel (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rb_device.h"

namespace rb {

constexpr int FUSED_MAX_UPW = 8;   // output row blocks (32 rows) per wave
constexpr int FUSED_GREG = 8;      // groups integrated per pass over the accumulators
constexpr int FUSED_MAX_REG_OPTS = 256;  // options whose sums live in registers

struct FusedUnit {
    float b;         // W1perp weight of the first-cell term
    float sc;        // sqrt(2H) N^-H
    int frag_off;    // offset of this unit's A-fragment table, in 16-byte entries
    int p_off;       // offset of the Np-entry table P_j = (j - off)^(2H), in floats
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
    int pad;
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

// per (path, payoff group) values shared by all strikes
struct PathPG {
    double lnS;      // ln S' = -rho^2 Iv / 2 + rho Is
    double sig;      // sqrt((1 - rho^2) Iv)
    float S;
    float invs;
    float lnSf;
    float sigf;
};

struct FusedArgs {
    int N, Np, NK, PB, TP;       // steps, padded steps (x32), 16-step k-blocks, path blocks, paths per tile
    int nunits, gmax, pgmax, out_size, gchunk, acc_in_lds;
    long M, ntiles, nitems;
    int tiles_per_item;
    PhiloxKey key;
    const FusedUnit* units;
    const FusedGroup* groups;
    const FusedPG* pgs;
    const FusedOption* opts;
    const uint4* frag;           // A fragments: per unit [Np/16][hi, lo][64 lanes] x 16 bytes
    const float* ptab;           // per unit Np floats
    double* partial;             // [out_size][nitems][2]
    int wave_pb[4];
    int wave_nu[4];              // row blocks of each wave, listed in decreasing order
    int wave_rb[4][FUSED_MAX_UPW];
};

size_t fused_lds_bytes(const FusedArgs& a);
// mode: 0 = FP16 hi+lo split, 1 = BF16 hi+lo split
hipError_t launch_fused(int mode, int upw, const FusedArgs& a, int grid, hipStream_t st);

}  // namespace rb
