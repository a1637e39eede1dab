// rb_fused_kernel.h -- the fused kernel template, instantiated by rb_fused.hip (Romano-Touzi
// estimator) and rb_fused_mc.hip (plain Monte Carlo); see rb_fused.hip for the stages.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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

// FP32 operands: the matrix product runs on v_mfma_f32_32x32x2_f32 (K = 2 per instruction).  The
// fragment tables and the LDS W1 tile keep the byte layout of the 16-bit modes: a lane's eight
// k-slots of a 16-wide chunk, k = 8 (l/32) + j, are eight floats, j = 0..3 in the first 16-byte
// half (where the 16-bit modes keep hi) and j = 4..7 in the second (lo).  Instruction j pairs the
// A value of slot j with the B value of slot j, as the 16-bit modes pair their k-slots.
template <>
struct SplitT<2> {
    typedef float T;
    typedef float V __attribute__((ext_vector_type(4)));
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ floatx16_t mma1(float a, float b, floatx16_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
};

// a float4 from global memory (16-byte aligned); the explicit address space gives a global load
// where the pointer's provenance is not visible to the compiler
__device__ __forceinline__ float4 load_global_f4(const float* p) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = *(const __attribute__((address_space(1))) f4*)(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// 16 bytes from global memory (see load_global_f)
__device__ __forceinline__ uint4 load_global_u4(const uint4* p) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 v = *(const __attribute__((address_space(1))) u4*)(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// a float from global memory (a generic pointer would give a FLAT load, after which the compiler
// waits for every outstanding vector memory load before any use of one)
__device__ __forceinline__ float load_global_f(const float* p) {
    return *(const __attribute__((address_space(1))) float*)(p);
}

// a wave-uniform pointer into the constant address space: the request's read-only tables read at
// wave-uniform indices then go through scalar loads (the scalar cache, no vector registers, no
// wait on the vector memory counter)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* uniform_const(const T* p) {
    // readfirstlane returns int: each half goes back through uint32_t before widening, or a low
    // half with bit 31 set would sign-extend into the high half
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(a))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(a >> 32))));
    return (const __attribute__((address_space(4))) T*)((uint64_t(hi) << 32) | uint64_t(lo));
}

// a record of read-only data at a wave-uniform address, through scalar loads
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword record");
    const auto q = uniform_const(reinterpret_cast<const uint32_t*>(p));
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); ++i) w[i] = q[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// a record of read-only data through global (not FLAT) vector loads, any address
template <class T>
__device__ __forceinline__ T load_global_rec(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword record");
    const __attribute__((address_space(1))) uint32_t* q = (const __attribute__((address_space(1))) uint32_t*)(p);
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); ++i) w[i] = q[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// 16 floats of a row table at p, p + 8, p + 16, p + 24 (four aligned float4: a lane's rows
// 8 c + 0..3, c = 0..3)
__device__ __forceinline__ void load_rows16(const float* p, float v[16]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float4 x = load_global_f4(p + 8 * c);
        v[4 * c + 0] = x.x;
        v[4 * c + 1] = x.y;
        v[4 * c + 2] = x.z;
        v[4 * c + 3] = x.w;
    }
}

// The W1 tile in LDS, per path block and 16-step chunk: 64 lane slots of 16 bytes as the MFMA B
// operand (slot 32 x + col holds steps 16 ks + 8 x + 0..7 of path column col), then, except in
// the FP16 mode, 64 more (fused_chunk_u4).  FP16 mode: W1 rounded to FP16.  BF16 mode: hi in the
// first 64, lo = RN(W1 - hi) in the second.  FP32 mode: steps 8 x + 0..3 in the first 64,
// 8 x + 4..7 in the second.

// W1 of path column `col`, step 16 ks + kk, from the LDS tile (the value the product used)
template <int MODE>
__device__ __forceinline__ float tile_w1(const uint4* chunk, int kk, int col) {
    if constexpr (MODE == 2) {
        return reinterpret_cast<const float*>(chunk + 64 * ((kk & 7) >> 2) + 32 * (kk >> 3) + col)[kk & 3];
    } else if constexpr (MODE == 0) {
        return float(reinterpret_cast<const _Float16*>(chunk + 32 * (kk >> 3) + col)[kk & 7]);
    } else {
        const uint4* Bp = chunk + 32 * (kk >> 3) + col;
        return float(reinterpret_cast<const __bf16*>(Bp)[kk & 7]) + float(reinterpret_cast<const __bf16*>(Bp + 64)[kk & 7]);
    }
}

// FP32 operands of path column `col`, steps 4 Q .. 4 Q + 3, from the LDS tile of path block
// `blk` (tile_w1 of the same steps): one 16-byte read.  Quad Q sits in chunk Q / 4, lane group
// (Q mod 4) / 2, half Q mod 2.
__device__ __forceinline__ void tile_w1_quad_f32(const uint4* Bf, int blk, int NK, int Q, int col, float w[4]) {
    const uint4* chunk = Bf + ((blk * NK + (Q >> 2)) * 2) * 64 + 32 * ((Q & 3) >> 1) + col;
    const float4 v = __builtin_bit_cast(float4, chunk[64 * (Q & 1)]);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
}

// one W1 element of the 16-bit modes from the dwords hi (and lo) of the tile: UPPER picks the
// upper 16 bits.  FP16: the FP16 value (lo unused).  BF16: hi + lo, the widening a shift.
template <int MODE, bool UPPER>
__device__ __forceinline__ float tile_elem(uint32_t hi, uint32_t lo) {
    if constexpr (MODE == 0) {
        return float(__builtin_bit_cast(_Float16, uint16_t(UPPER ? hi >> 16 : hi & 0xffffu)));
    } else {
        return UPPER ? __uint_as_float(hi & 0xffff0000u) + __uint_as_float(lo & 0xffff0000u)
                     : __uint_as_float(hi << 16) + __uint_as_float(lo << 16);
    }
}

__device__ __forceinline__ double half_sum_f64(double v) {
    // sum over the 32 lanes of one half-wave; every lane of the half receives the total
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// GR: groups integrated per pass over the accumulators (at most FUSED_GREG)
template <int MODE, int UPW, bool MC, int GR>
__global__ void __launch_bounds__(256, UPW >= 8 ? 1 : 2) k_fused(FusedBatchArgs B) {
    static_assert(GR <= FUSED_GREG, "pgmax is sized for windows of FUSED_GREG groups");
    constexpr int NS = MC ? 3 : 2;                        // sums per (path, group)
    typedef SplitT<MODE> S;
    typedef typename S::V V;
    typedef typename S::V4 V4;
    constexpr int CH = fused_chunk_u4(MODE);               // 16-byte units per W1 tile chunk
    extern __shared__ uint4 lds4[];
    __shared__ int s_item;
    __shared__ FusedReq s_req;   // the current item's request record
    static_assert(sizeof(FusedReq) % 4 == 0, "dword copy");
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, col = lane & 31;
    uint4* Bf = lds4;                                                   // [PB][NK][CH]
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
        // the request record, once per item, staged in LDS (reads through the global pointer
        // would be reloaded after every global store); the loop controls read from it are made
        // wave-uniform (scalar branches, no exec masks).  A per-lane copy of the record (kept in
        // scratch, since its per-wave tables are indexed by the wave number: 312 B written per
        // lane per item) was faster with four row blocks per wave in earlier builds; with the
        // branch-free integral loop the LDS record measured 8.70 vs 9.18 ms at N = 512, M = 2^22,
        // and the serving trace 2.23e11 vs 2.08e11 paths*steps/s.
        {
            const int* src = reinterpret_cast<const int*>(B.reqs + B.item_req[gi]);
            int* dst = reinterpret_cast<int*>(&s_req);
            for (int i = tid; i < int(sizeof(FusedReq) / 4); i += 256) dst[i] = src[i];
            __syncthreads();
        }
        const FusedReq& A = s_req;
        const long item = B.item_local[gi];
        const int NK = A.NK;
        const int pbw = __builtin_amdgcn_readfirstlane(A.wave_pb[wave]);
        const int nuw = __builtin_amdgcn_readfirstlane(A.wave_nu[wave]);
        double a1 = 0.0, a2 = 0.0;  // sums of option tid (register accumulation)
        if (A.acc_in_lds)
            for (int o = tid; o < 2 * A.out_size; o += 256) accopt[o] = 0.0;
        const bool big = item < A.items_big;
        const long t_begin = big ? item * A.tiles_per_item
                                 : A.items_big * A.tiles_per_item + (item - A.items_big) * A.tiles_small;
        const long t_len = big ? A.tiles_per_item : A.tiles_small;
        const long t_end = t_begin + t_len < A.ntiles ? t_begin + t_len : A.ntiles;
        for (long tile = t_begin; tile < t_end; ++tile) {
            const long p0 = tile * A.TP;
            // ---- 1. Gaussians W1 -> B fragments (FP16, BF16 hi + lo, or FP32).  Every quad of
            //         the padded steps is generated, and the paths past M keep their Gaussians:
            //         steps >= N reach only output rows >= N (the Toeplitz operand is zero past
            //         N - 1), whose terms the row term +inf zeroes, and the results of paths past
            //         M are dropped.  Four quads per trip, so four Philox chains interleave.
            const int nq = A.Np >> 2;
            {
                const int pl = tid & (A.TP - 1);   // TP is 32, 64 or 128
                const int pb = pl >> 5;
                const long p = p0 + pl;
                const int step = 256 >> A.log2TP;  // divides nq (a multiple of 8)
                auto put = [&](int q, const float z[4]) {
                    const int ks = q >> 2, hh = (q & 3) >> 1, j0 = (q & 1);
                    const int ln = 32 * hh + (pl & 31);
                    uint4* base = Bf + (pb * NK + ks) * CH;
                    if constexpr (MODE == 2) {
                        const V4 v = {z[0], z[1], z[2], z[3]};
                        base[64 * j0 + ln] = __builtin_bit_cast(uint4, v);
                    } else if constexpr (MODE == 0) {
                        typedef _Float16 h4 __attribute__((ext_vector_type(4)));
                        const h4 hv = {(_Float16)z[0], (_Float16)z[1], (_Float16)z[2], (_Float16)z[3]};
                        reinterpret_cast<uint2*>(base + ln)[j0] = __builtin_bit_cast(uint2, hv);
                    } else {
                        V4 hi, lo;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            hi[j] = S::cvt(z[j]);
                            lo[j] = S::cvt(z[j] - float(hi[j]));
                        }
                        reinterpret_cast<uint2*>(base + ln)[j0] = __builtin_bit_cast(uint2, hi);
                        reinterpret_cast<uint2*>(base + 64 + ln)[j0] = __builtin_bit_cast(uint2, lo);
                    }
                };
                int q = tid >> A.log2TP;
                for (; q + 3 * step < nq; q += 4 * step) {
                    float z[4][4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) normals4_f32(A.key, uint64_t(p), 0u, uint32_t(q + v * step), z[v]);
#pragma unroll
                    for (int v = 0; v < 4; ++v) put(q + v * step, z[v]);
                }
                for (; q < nq; q += step) {
                    float z[4];
                    normals4_f32(A.key, uint64_t(p), 0u, uint32_t(q), z);
                    put(q, z);
                }
            }
            __syncthreads();

            for (int u = 0; u < A.nunits; ++u) {
                const FusedUnit U = load_uniform(A.units + u);
                // ---- 2. Toeplitz GEMM on the matrix cores
                floatx16_t acc[UPW];
#pragma unroll
                for (int s = 0; s < UPW; ++s)
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[s][r] = 0.f;
                if (pbw >= 0) {
                    const uint4* F = A.frag + U.frag_off;
                    int rbs[UPW];   // the wave's row blocks, decreasing (wave-uniform)
#pragma unroll
                    for (int s = 0; s < UPW; ++s) rbs[s] = __builtin_amdgcn_readfirstlane(A.wave_rb[wave][s]);
                    // d runs down from 32 rbs[0] to -16 in steps of 16 across the phases; every
                    // phase has an even number of steps (2 (rbs[k-1] - rbs[k]), or 2 rbs[k-1] + 2
                    // in the last), so the steps go in pairs d, d - 16 with the A fragments in
                    // two register sets X and Y: a set is reloaded for d - 32 (d - 48) as soon as
                    // its products are issued, and the load has the other set's products to
                    // arrive (global loads: the wait before a product counts only the newer
                    // ones).
                    auto fa = [&](int dd, int half) {
                        const int dc = dd > -16 ? dd : -16;
                        return __builtin_bit_cast(V, load_global_u4(F + ((dc + 16) >> 4) * 128 + 64 * half + lane));
                    };
                    auto step = [&](int dd, const V& ahi, const V& alo, auto kc) {
                        constexpr int k = decltype(kc)::value;
                        V bhi[UPW], blo[UPW];
#pragma unroll
                        for (int s = 0; s < k; ++s) {
                            const int ks = (32 * rbs[s] - dd) >> 4;
                            const uint4* Bp = Bf + (pbw * NK + ks) * CH;
                            bhi[s] = __builtin_bit_cast(V, Bp[lane]);
                            if constexpr (MODE != 0) blo[s] = __builtin_bit_cast(V, Bp[64 + lane]);
                        }
                        if constexpr (MODE == 2) {
                            // k-slots 0..3 from the first halves, 4..7 from the second
#pragma unroll
                            for (int j = 0; j < 4; ++j)
#pragma unroll
                                for (int s = 0; s < k; ++s) acc[s] = S::mma1(ahi[j], bhi[s][j], acc[s]);
#pragma unroll
                            for (int j = 0; j < 4; ++j)
#pragma unroll
                                for (int s = 0; s < k; ++s) acc[s] = S::mma1(alo[j], blo[s][j], acc[s]);
                        } else {
                            // A hi W1 (+ A hi W1lo, BF16) + A lo W1
#pragma unroll
                            for (int s = 0; s < k; ++s) acc[s] = S::mma(ahi, bhi[s], acc[s]);
                            if constexpr (MODE == 1) {
#pragma unroll
                                for (int s = 0; s < k; ++s) acc[s] = S::mma(ahi, blo[s], acc[s]);
                            }
#pragma unroll
                            for (int s = 0; s < k; ++s) acc[s] = S::mma(alo, bhi[s], acc[s]);
                        }
                    };
                    int d = 32 * rbs[0];
                    V xhi = fa(d, 0), xlo = fa(d, 1), yhi = fa(d - 16, 0), ylo = fa(d - 16, 1);
                    // phase k: row blocks 0..k-1 of the wave are active (blocks sorted descending)
                    auto phase = [&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        // the last phase ends with the pair (0, -16), which reloads nothing (a
                        // load still in flight past the products would hold up the next stage)
                        const bool last = !(k < nuw);
                        const int dlo = last ? 16 : 32 * rbs[k < UPW ? k : 0] + 16;
                        for (; d >= dlo; d -= 32) {
                            // the scheduling barriers keep each reload right after its set's
                            // products instead of sinking it to the end of the pair
                            step(d, xhi, xlo, kc);
                            xhi = fa(d - 32, 0);
                            xlo = fa(d - 32, 1);
                            __builtin_amdgcn_sched_barrier(0);
                            step(d - 16, yhi, ylo, kc);
                            yhi = fa(d - 48, 0);
                            ylo = fa(d - 48, 1);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        if (last) {
                            step(0, xhi, xlo, kc);
                            step(-16, yhi, ylo, kc);
                        }
                    };
                    phase(std::integral_constant<int, 1>());
                    if constexpr (UPW >= 2) if (nuw >= 2) phase(std::integral_constant<int, 2>());
                    if constexpr (UPW >= 3) if (nuw >= 3) phase(std::integral_constant<int, 3>());
                    if constexpr (UPW >= 4) if (nuw >= 4) phase(std::integral_constant<int, 4>());
                    if constexpr (UPW >= 5) if (nuw >= 5) phase(std::integral_constant<int, 5>());
                    if constexpr (UPW >= 6) if (nuw >= 6) phase(std::integral_constant<int, 6>());
                    if constexpr (UPW >= 7) if (nuw >= 7) phase(std::integral_constant<int, 7>());
                    if constexpr (UPW >= 8) if (nuw >= 8) phase(std::integral_constant<int, 8>());
                    // ---- 3a. Wtilde = sc (Y' + b W1perp), W1perp from the same counters (stream 1)
                    const long p = p0 + 32 * pbw + col;
#pragma unroll
                    for (int s = 0; s < UPW; ++s) {
                        if (s >= nuw) break;
                        const int rb = A.wave_rb[wave][s];
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            float z[4];  // paths past M too: their results are dropped
                            normals4_f32(A.key, uint64_t(p), 1u, uint32_t(8 * rb + 2 * c + h), z);
#pragma unroll
                            for (int t = 0; t < 4; ++t)
                                acc[s][4 * c + t] = U.sc * (acc[s][4 * c + t] + U.b * z[t]);
                        }
                    }
                }
                // ---- 3b. integrals of every group of the unit, GR groups per pass
                for (int g0 = U.g_begin; g0 < U.g_end; g0 += GR) {
                    const int gn_ = U.g_end - g0 < GR ? U.g_end - g0 : GR;
                    const int gn = __builtin_amdgcn_readfirstlane(gn_);
                    float c1[GR];
                    double sv[GR], ss[GR], sp[GR];
                    // FP32 partial sums: per 16-row block (Romano-Touzi), or over all row blocks of
                    // the wave (plain Monte Carlo, at most 128 terms a lane; frees the registers
                    // for the third sum)
                    float fv[GR], fs[GR], fp[GR];
#pragma unroll
                    for (int gg = 0; gg < GR; ++gg) {
                        sv[gg] = ss[gg] = sp[gg] = 0.0;
                        fv[gg] = fs[gg] = fp[gg] = 0.f;
                        c1[gg] = gg < gn ? uniform_const(A.groups)[g0 + gg].c1 : 0.f;
                    }
                    if (pbw >= 0) {
                        // row terms al P_j + be: group g0 + gg at gg Np + j - 1
                        const float* Qt = A.ptab + U.p_off + (g0 - U.g_begin) * A.Np;
                        const long p = p0 + 32 * pbw + col;
#pragma unroll
                        for (int s = 0; s < UPW; ++s) {
                            if (s >= nuw) break;
                            const int rb = __builtin_amdgcn_readfirstlane(A.wave_rb[wave][s]);
                            if constexpr (!MC) {
#pragma unroll
                                for (int gg = 0; gg < GR; ++gg) fv[gg] = fs[gg] = 0.f;
                            }
                            // plain Monte Carlo: Wperp_j (stream 2) at the rows j = row + 1 this lane
                            // integrates.  Quad 8 rb + 2 c + h holds steps 4 (8 rb + 2 c + h) + 0..3,
                            // so t = 0..2 come from the lane's own quad and t = 3 from element 0 of
                            // the next quad: the other half-wave's quad (c, or c + 1 for h = 1), and
                            // for (c = 3, h = 1) the first quad of the next row block.
                            float wq[16];
                            if constexpr (MC) {
                                float zc[4][4], ze[4];
#pragma unroll
                                for (int c = 0; c < 4; ++c)
                                    normals4_f32(A.key, uint64_t(p), 2u, uint32_t(8 * rb + 2 * c + h), zc[c]);
                                normals4_f32(A.key, uint64_t(p), 2u, uint32_t(8 * rb + 8), ze);
                                float y[4];
#pragma unroll
                                for (int c = 0; c < 4; ++c) y[c] = __shfl_xor(zc[c][0], 32, 64);
#pragma unroll
                                for (int c = 0; c < 4; ++c) {
#pragma unroll
                                    for (int t = 0; t < 3; ++t) wq[4 * c + t] = zc[c][t + 1];
                                    wq[4 * c + 3] = h == 0 ? y[c] : (c < 3 ? y[c < 3 ? c + 1 : 3] : ze[0]);
                                }
                            }
                            // W1_j of the lane's rows j = 32 rb + 8 c + 4 h + t + 1, t = 0..3, read
                            // ahead of the row loop: elements 1..3 of quad 8 rb + 2 c + h and
                            // element 0 of the next quad (which lies past the tile only for row
                            // j = Np, whose row term is +inf; that read stays on the last quad)
                            float wv[16];
                            const float* Qr = Qt + 32 * rb + 4 * h;
                            if constexpr (MODE == 2) {
#pragma unroll
                                for (int c = 0; c < 4; ++c) {
                                    const int Q = 8 * rb + 2 * c + h;
                                    float qa[4], qb[4];
                                    tile_w1_quad_f32(Bf, pbw, NK, Q, col, qa);
                                    tile_w1_quad_f32(Bf, pbw, NK, Q + 1 < 4 * NK ? Q + 1 : Q, col, qb);
                                    wv[4 * c + 0] = qa[1];
                                    wv[4 * c + 1] = qa[2];
                                    wv[4 * c + 2] = qa[3];
                                    wv[4 * c + 3] = qb[0];
                                }
                            } else {
                                // 16-bit modes, byte offsets from slot col of chunk 2 rb (chunks of
                                // 16 CH bytes): quad Q is the uint2 h of slot 32 (c & 1) + col in
                                // chunk 2 rb + c / 2 (BF16 lo 1024 B further).  Element 0 of quad
                                // Q + 1 begins the next uint2 (h = 0), slot 32 + col (h = 1, c
                                // even) or slot col of the next chunk (h = 1, c odd); for c = 3 in
                                // the last row block that is past the tile, and the row (j = Np)
                                // has P = +inf, so the read takes element 0 of the lane's own slot
                                const char* Wr = reinterpret_cast<const char*>(Bf + (pbw * NK + 2 * rb) * CH + col);
                                const int nb_even = h ? 512 : 8, nb_odd = h ? 16 * CH - 512 : 8;
                                const int nb3 = 2 * rb + 2 < NK ? nb_odd : (h ? 0 : 8);
#pragma unroll
                                for (int c = 0; c < 4; ++c) {
                                    const char* slot = Wr + (c >> 1) * (16 * CH) + (c & 1) * 512;
                                    const int nb = c == 3 ? nb3 : (c & 1) ? nb_odd : nb_even;
                                    const uint2 ha = *reinterpret_cast<const uint2*>(slot + 8 * h);
                                    const uint32_t hb = *reinterpret_cast<const uint32_t*>(slot + nb);
                                    uint2 la = make_uint2(0u, 0u);
                                    uint32_t lb = 0u;
                                    if constexpr (MODE == 1) {
                                        la = *reinterpret_cast<const uint2*>(slot + 8 * h + 1024);
                                        lb = *reinterpret_cast<const uint32_t*>(slot + nb + 1024);
                                    }
                                    wv[4 * c + 0] = tile_elem<MODE, true>(ha.x, la.x);
                                    wv[4 * c + 1] = tile_elem<MODE, false>(ha.y, la.y);
                                    wv[4 * c + 2] = tile_elem<MODE, true>(ha.y, la.y);
                                    wv[4 * c + 3] = tile_elem<MODE, false>(hb, lb);
                                }
                            }
                            // rows j = 32 rb + row + 1 >= N have the row term +inf (P = +inf,
                            // al > 0), so rr = 0 and they add nothing: no per-row test.  Groups
                            // outside, rows inside, and no branch at all: every group slot of the
                            // window is integrated, so the 16 GR rows are one basic block in which
                            // the compiler interleaves the groups' dependent sums.  A slot past gn
                            // has c1 = 0 and the last group's row terms, so its sums are finite,
                            // and they are never read.  Each group's sum still runs over
                            // r = 0..15 in order.  (One uniform branch per group, so that dead
                            // slots cost nothing: N = 512, M = 2^22, five maturities 6.55 against
                            // 6.43 ms, plain MC 6.85 against 6.62, serving trace 3.16e11 against
                            // 3.19e11 paths*steps/s; both forms in one kernel spilled more and
                            // lost on the trace.)  The
                            // lane's 16 row terms of a group are four aligned float4 loads, those
                            // of the next group issued before this group's rows (two buffers by
                            // the parity of the group); slots past gn reload the last group's.
                            float qv[2][16];
                            load_rows16(Qr, qv[0]);
#pragma unroll
                            for (int gg = 0; gg < GR; ++gg) {
                                if (gg + 1 < GR)
                                    load_rows16(Qr + (gg + 1 < gn ? gg + 1 : gn - 1) * A.Np, qv[(gg + 1) & 1]);
#pragma unroll
                                for (int r = 0; r < 16; ++r) {
                                    const float rr = __builtin_amdgcn_exp2f(fmaf(acc[s][r], c1[gg], -qv[gg & 1][r]));
                                    fv[gg] = fmaf(rr, rr, fv[gg]);
                                    fs[gg] = fmaf(rr, wv[r], fs[gg]);
                                    if constexpr (MC) fp[gg] = fmaf(rr, wq[r], fp[gg]);
                                }
                            }
                            if constexpr (!MC) {
#pragma unroll
                                for (int gg = 0; gg < GR; ++gg) {
                                    sv[gg] += double(fv[gg]);
                                    ss[gg] += double(fs[gg]);
                                }
                            }
                        }
                    }
                    if constexpr (MC) {
#pragma unroll
                        for (int gg = 0; gg < GR; ++gg) {
                            sv[gg] = double(fv[gg]);
                            ss[gg] = double(fs[gg]);
                            sp[gg] = double(fp[gg]);
                        }
                    }
#pragma unroll
                    for (int gg = 0; gg < GR; ++gg) {
                        sv[gg] += __shfl_xor(sv[gg], 32, 64);
                        ss[gg] += __shfl_xor(ss[gg], 32, 64);
                        if constexpr (MC) sp[gg] += __shfl_xor(sp[gg], 32, 64);
                    }
                    // cross-wave sums in chunks of gchunk groups (LDS budget)
                    const int pg_lo = uniform_const(A.groups)[g0].pg_begin;
                    for (int c0 = 0; c0 < gn; c0 += A.gchunk) {
                        const int cn = gn - c0 < A.gchunk ? gn - c0 : A.gchunk;
                        if (h == 0) {
#pragma unroll
                            for (int gg = 0; gg < GR; ++gg)
                                if (gg >= c0 && gg < c0 + cn) {
                                    double* w = wpart + ((wave * A.gchunk + (gg - c0)) * 32 + col) * NS;
                                    w[0] = sv[gg];
                                    w[1] = ss[gg];
                                    if constexpr (MC) w[2] = sp[gg];
                                }
                        }
                        __syncthreads();
                        for (int e = tid; e < A.TP * cn; e += 256) {
                            // FP64 record of (path, payoff group): no multiply-add contraction,
                            // so every instantiation (groups per pass, row blocks per wave)
                            // computes the same bits and a request prices identically alone or
                            // in any batch
#pragma clang fp contract(off)
                            const int pl = e & (A.TP - 1), gg = e >> A.log2TP;
                            const int pb = pl >> 5, c = pl & 31;
                            const FusedGroup G = load_global_rec(A.groups + g0 + c0 + gg);
                            const double w10 = double(tile_w1<MODE>(Bf + (pb * NK) * CH, 0, c));  // k = 0
                            double a = G.xi, b = G.sxi * w10;   // j = 0: v_0 = xi
                            double cp = 0.0;
                            if constexpr (MC) {
                                float z0[4];
                                normals4_f32(A.key, uint64_t(p0 + pl), 2u, 0u, z0);
                                cp = G.sxi * double(z0[0]);
                            }
                            for (int w = 0; w < 4; ++w)
                                if (A.wave_pb[w] == pb) {
                                    a += wpart[((w * A.gchunk + gg) * 32 + c) * NS + 0];
                                    b += wpart[((w * A.gchunk + gg) * 32 + c) * NS + 1];
                                    if constexpr (MC) cp += wpart[((w * A.gchunk + gg) * 32 + c) * NS + 2];
                                }
                            const double Iv = G.dt * a, Is = G.sdt * b;
                            if constexpr (MC) {
                                // ln S_T = rho Is + sqrt(1 - rho^2) Ip - Iv / 2 (log-Euler, left point)
                                const double Ip = G.sdt * cp;
                                for (int pg = G.pg_begin; pg < G.pg_end; ++pg) {
                                    const double rho = load_global_rec(A.pgs + pg).rho;
                                    PathPG q;
                                    q.lnS = rho * Is + sqrt(1.0 - rho * rho) * Ip - 0.5 * Iv;
                                    q.sig = exp(q.lnS);
                                    q.S = float(q.sig);
                                    q.ik = q.A = q.sg = 0.f;
                                    pgd[pl * A.pgmax + (pg - pg_lo)] = q;
                                }
                                continue;
                            }
                            for (int pg = G.pg_begin; pg < G.pg_end; ++pg) {
                                const double rho = load_global_rec(A.pgs + pg).rho;
                                PathPG q;
                                q.sig = sqrt((1.0 - rho * rho) * Iv);
                                q.lnS = -0.5 * rho * rho * Iv + rho * Is;
                                pathpg_f32(q);
                                pgd[pl * A.pgmax + (pg - pg_lo)] = q;
                            }
                        }
                        __syncthreads();
                    }
                    // ---- 4. payoffs of the options of groups g0 .. g0+gn-1.  Register mode: the
                    //         S = 2^log2slices lanes t = o*S .. o*S+S-1 share option o, each
                    //         summing every S-th path; a fixed butterfly leaves the total in lane
                    //         o*S, which owns option o's FP64 sums.  With one lane per option,
                    //         options 0 .. pay2-1 take lane pairs (threads 0 .. 2 pay2 - 1, whole
                    //         waves) and option o >= pay2 thread o + pay2 (t = tid - pay2).  LDS
                    //         mode (> 256 options): one thread per option, sums in LDS.
                    const int o_lo = uniform_const(A.groups)[g0].o_begin, o_hi = uniform_const(A.groups)[g0 + gn - 1].o_end;
                    const int np = A.M - p0 < A.TP ? int(A.M - p0) : A.TP;
                    if (A.acc_in_lds) {
                        for (int oi = o_lo + tid; oi < o_hi; oi += 256) {
                            const FusedOption O = load_global_rec(A.opts + oi);
                            const int pgl = O.pg - pg_lo;
                            const OptF of = opt_f32(O.lnK, O.K);
                            double s1 = 0.0, s2 = 0.0;
                            for (int pl = 0; pl < np; ++pl) {
                                const double x = MC ? payoff_mc(pgd[pl * A.pgmax + pgl], O.K)
                                                    : payoff_mixed(pgd[pl * A.pgmax + pgl], of, O.lnK, O.K);
                                s1 += x;
                                s2 += x * x;
                            }
                            accopt[2 * oi + 0] += s1;
                            accopt[2 * oi + 1] += s2;
                        }
                    } else {
                        const bool pair = tid < 2 * A.pay2;                   // wave-uniform
                        const int lg = pair ? 1 : A.log2slices;
                        const int t = pair ? tid : tid - 2 * A.pay2 + (A.pay2 << A.log2slices);
                        const int S = 1 << lg;
                        const int oi = t >> lg, sl = t & (S - 1);
                        if (oi >= o_lo && oi < o_hi) {
                            const FusedOption O = load_global_rec(A.opts + oi);
                            const int pgl = O.pg - pg_lo;
                            const OptF of = opt_f32(O.lnK, O.K);
                            double s1 = 0.0, s2 = 0.0;
                            for (int pl = sl; pl < np; pl += S) {
                                const double x = MC ? payoff_mc(pgd[pl * A.pgmax + pgl], O.K)
                                                    : payoff_mixed(pgd[pl * A.pgmax + pgl], of, O.lnK, O.K);
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
                    // no barrier here: the W1 tile was last read before the record barrier
                    // above, so the next tile's Gaussians may overwrite it while other waves
                    // still price; the records are rewritten, and the partials reread, only
                    // after the next pass's or tile's barriers, which every wave reaches after
                    // its payoffs
                }
            }
        }
        // ---- item partial sums (indexed by the output index I)
        if (A.acc_in_lds) {
            __syncthreads();   // option o's LDS sums were added by thread (o - o_lo) mod 256
            for (int o = tid; o < A.out_size; o += 256) {
                const int I = A.opts[o].out;
                A.partial[(long(I) * A.nitems + item) * 2 + 0] = accopt[2 * o + 0];
                A.partial[(long(I) * A.nitems + item) * 2 + 1] = accopt[2 * o + 1];
            }
        } else {
            const bool pair = tid < 2 * A.pay2;   // the payoff stage's lanes
            const int lg = pair ? 1 : A.log2slices;
            const int t = pair ? tid : tid - 2 * A.pay2 + (A.pay2 << A.log2slices);
            if ((t & ((1 << lg) - 1)) == 0 && (t >> lg) < A.out_size) {
                const int I = A.opts[t >> lg].out;
                A.partial[(long(I) * A.nitems + item) * 2 + 0] = a1;
                A.partial[(long(I) * A.nitems + item) * 2 + 1] = a2;
            }
        }
        __syncthreads();
    }
}


size_t fused_lds_bytes(const FusedBatchArgs& b);

template <int MODE, int UPW, bool MC, int GR>
static hipError_t launch_t(const FusedBatchArgs& a, int grid, hipStream_t st) {
    const size_t lds = fused_lds_bytes(a);
    if (lds > 65536) {
        const hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_fused<MODE, UPW, MC, GR>), lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_fused<MODE, UPW, MC, GR>), dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
}

template <bool MC, int GR>
static hipError_t launch_mu(int mode, int upw, const FusedBatchArgs& a, int grid, hipStream_t st) {
    switch (mode * 16 + upw) {
        case 1: return launch_t<0, 1, MC, GR>(a, grid, st);
        case 2: return launch_t<0, 2, MC, GR>(a, grid, st);
        case 4: return launch_t<0, 4, MC, GR>(a, grid, st);
        case 8: return launch_t<0, 8, MC, GR>(a, grid, st);
        case 17: return launch_t<1, 1, MC, GR>(a, grid, st);
        case 18: return launch_t<1, 2, MC, GR>(a, grid, st);
        case 20: return launch_t<1, 4, MC, GR>(a, grid, st);
        case 24: return launch_t<1, 8, MC, GR>(a, grid, st);
        case 33: return launch_t<2, 1, MC, GR>(a, grid, st);
        case 34: return launch_t<2, 2, MC, GR>(a, grid, st);
        case 36: return launch_t<2, 4, MC, GR>(a, grid, st);
        case 40: return launch_t<2, 8, MC, GR>(a, grid, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rb
