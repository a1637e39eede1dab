// rb_device.h -- device helpers shared by the rough-Bergomi kernels (gfx950).
//
// Counter-based Gaussians: normal j of stream s of path p is Box-Muller on
// Philox4x32-10(counter = {j/4, p_lo, p_hi, s}, key = fold(seed)).  Any (path, step) can
// be regenerated in any thread, so results do not depend on launch shape.  This replaces
// the per-thread mt19937_64 streams of the reference (RNG::RNG, rBergLib/rBergomiMT.cpp:63-71)
// whose output depends on the thread count.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rb {

struct PhiloxKey {
    uint32_t k0, k1;
};

// a ^ b ^ c: on the device one v_bitop3_b32 (truth table 0x96), which the compiler does not form
// from two XORs by itself; the host pass (probes, tables) uses the operators
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11): multipliers 0xD2511F53,
// 0xCD9E8D57, Weyl key increments 0x9E3779B9, 0xBB67AE85.
__host__ __device__ __forceinline__ uint4 philox10(uint4 c, PhiloxKey key) {
    uint32_t k0 = key.k0, k1 = key.k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        uint4 n;
        n.x = xor3(uint32_t(p1 >> 32), c.y, k0);
        n.y = uint32_t(p1);
        n.z = xor3(uint32_t(p0 >> 32), c.w, k1);
        n.w = uint32_t(p0);
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__host__ __device__ __forceinline__ uint4 philox_block(PhiloxKey key, uint64_t path, uint32_t stream,
                                              uint32_t q) {
    return philox10(make_uint4(q, uint32_t(path), uint32_t(path >> 32), stream), key);
}

// Box-Muller in FP64 on u = (x + 1/2) 2^-32: (r cos 2 pi u1, r sin 2 pi u1), r = sqrt(-2 ln u0)
__device__ __forceinline__ void box_muller_f64(uint32_t x0, uint32_t x1, double& z0, double& z1) {
    const double u0 = (double(x0) + 0.5) * 2.3283064365386963e-10;
    const double u1 = (double(x1) + 0.5) * 2.3283064365386963e-10;
    const double r = sqrt(-2.0 * log(u0));
    double s, c;
    sincospi(2.0 * u1, &s, &c);
    z0 = r * c;
    z1 = r * s;
}

// The same transform in FP32 (mixed-precision kernels): identical counters, FP32 rounding.
// Written with the hardware instructions directly, so that every call site compiles to the same
// instruction sequence (a library sqrt was lowered correctly rounded at some call sites and as a
// scaled one-ulp square root at others, so the same counter gave different FP32 normals in
// different kernel instantiations).  u0 >= 2^-33 is normal; -2 ln u0 is 0 or >= 1e-7, never
// subnormal, so the unscaled v_sqrt_f32 (1 ulp) applies.  The hardware sine and cosine take their
// argument in turns: sin(2 pi u1) = v_sin_f32(u1).  u = (x + 1/2) 2^-32 as one FMA,
// x 2^-32 + 2^-33: the product is exact, so it rounds once where (x + 0.5) rounds once, and the
// two agree for every 32-bit x (checked exhaustively on the host).
__device__ __forceinline__ void box_muller_f32(uint32_t x0, uint32_t x1, float& z0, float& z1) {
    // the normals are rounded products: no contraction into a caller's add (the BF16 split
    // z - hi would otherwise become fma(r, cos, -hi) in some instantiations and not in others)
#pragma clang fp contract(off)
    const float u0 = fmaf(float(x0), 2.3283064365386963e-10f, 1.1641532182693481e-10f);
    const float u1 = fmaf(float(x1), 2.3283064365386963e-10f, 1.1641532182693481e-10f);
    // ln u0 = log2(u0) * ln 2
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u0));
    z0 = r * __builtin_amdgcn_cosf(u1);
    z1 = r * __builtin_amdgcn_sinf(u1);
}

// Four normals of path p, stream s, steps 4q..4q+3 in FP64
__device__ __forceinline__ void normals4_f64(PhiloxKey key, uint64_t p, uint32_t s, uint32_t q,
                                             double z[4]) {
    const uint4 x = philox_block(key, p, s, q);
    box_muller_f64(x.x, x.y, z[0], z[1]);
    box_muller_f64(x.z, x.w, z[2], z[3]);
}

__device__ __forceinline__ void normals4_f32(PhiloxKey key, uint64_t p, uint32_t s, uint32_t q,
                                             float z[4]) {
    const uint4 x = philox_block(key, p, s, q);
    box_muller_f32(x.x, x.y, z[0], z[1]);
    box_muller_f32(x.z, x.w, z[2], z[3]);
}

// Phi(x) = erfc(-x/sqrt 2)/2, rBergLib/BlackScholes.cpp:16-19
__device__ __forceinline__ double phi_f64(double x) { return 0.5 * erfc(-x * 0.7071067811865476); }

// Inverse standard normal CDF in FP64 for 0 < u < 1: the rational start of Abramowitz & Stegun
// 26.2.23 (|error| < 4.5e-4) on q = min(u, 1 - u), then three Halley steps on Phi(x) - q (cubic
// convergence: 4.5e-4 -> 1e-10 -> machine precision).  Used for Sobol' points as normalQMC_sample
// does with Phi_inv (rBergomi/qmc.cpp:26-41).
__device__ inline double norm_inv_f64(double u) {
    const double q = u < 0.5 ? u : 1.0 - u;
    const double t = sqrt(-2.0 * log(q));
    double x = -(t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                         (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308))));
    for (int it = 0; it < 3; ++it) {
        const double f = phi_f64(x) - q;
        const double d = 0.3989422804014327 * exp(-0.5 * x * x);
        const double r = f / d;
        x -= r / (1.0 + 0.5 * x * r);
    }
    return u < 0.5 ? x : -x;
}

// Romano-Touzi payoff: C_BS(S', K, tau = 1, sigma) with sigma = sqrt((1-rho^2) Iv),
// ln S' = -rho^2 Iv / 2 + rho Is  (rBergLib/rBergomiMT.cpp:383-388, BlackScholes.cpp:22-26).
// ln(S'/K) is formed as lnS - lnK, which avoids a log per strike.
__device__ __forceinline__ double rt_payoff_f64(double Iv, double Is, double rho, double K,
                                                double lnK) {
    const double sig = sqrt((1.0 - rho * rho) * Iv);
    const double lnS = -0.5 * rho * rho * Iv + rho * Is;
    const double S = exp(lnS);
    const double d1 = (lnS - lnK + 0.5 * sig * sig) / sig;
    const double d2 = d1 - sig;
    return phi_f64(d1) * S - phi_f64(d2) * K;
}

__device__ __forceinline__ double bs_price_f64(double K, double tau, double sig, bool call) {
    const double st = sig * sqrt(tau);
    const double d1 = (-log(K) + 0.5 * st * st) / st;
    const double d2 = d1 - st;
    return call ? phi_f64(d1) - K * phi_f64(d2) : -phi_f64(-d1) + K * phi_f64(-d2);
}

// Implied vol of a call with S0 = 1, r = 0.  In-the-money calls are solved on the put
// (put-call parity, rBergLib/BlackScholes.cpp:36-40).  The reference brackets with Boost
// to a 2^-10 tolerance; here a bracketed Newton iteration runs to |dsigma| <= 1e-14 sigma.
// NaN when the price lies outside the no-arbitrage range.
__device__ inline double implied_vol_f64(double price, double K, double tau) {
    bool call = true;
    double target = price;
    if (1.0 > K) {
        call = false;
        target = price - 1.0 + K;
    }
    const double upper = call ? 1.0 : K;
    if (!(target > 0.0) || !(target < upper) || !(tau > 0.0)) return __builtin_nan("");
    double lo = 0.0, hi = 1.0;
    while (bs_price_f64(K, tau, hi, call) < target) {
        lo = hi;
        hi *= 2.0;
        if (hi > 1e6) return __builtin_nan("");
    }
    double s = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
        const double f = bs_price_f64(K, tau, s, call) - target;
        if (f > 0) hi = s; else lo = s;
        const double d1 = (-log(K) + 0.5 * s * s * tau) / (s * sqrt(tau));
        const double vega = exp(-0.5 * d1 * d1) * 0.3989422804014327 * sqrt(tau);
        double sn = vega > 0 ? s - f / vega : 0.5 * (lo + hi);
        if (!(sn > lo && sn < hi)) sn = 0.5 * (lo + hi);
        const double ds = fabs(sn - s);
        s = sn;
        if (ds <= 1e-14 * s || hi - lo <= 1e-15 * s) break;
    }
    return s;
}

// The reference's implied vol to the reference's tolerance (rBergLib/BlackScholes.cpp:36-92):
// in-the-money calls (K < 1) are solved on the put through put-call parity, as IV_call does.  Its
// solver is Boost.Math's bracket_and_solve_root on f(sigma) = target - BS(sigma), which decreases
// in sigma (is_rising = false): from the guess 0.2 the bracket walks by the factor 2 -- down while
// f < 0, up while f > 0 -- until f changes sign (at most maxit = 20 evaluations), then the bracket
// is narrowed until eps_tolerance(11) holds, |b - a| <= 2^-10 min(a, b), and its midpoint is
// returned.  Boost narrows with TOMS 748 (Alefeld, Potra and Shi, ACM TOMS 21(3), 1995); here the
// bracket is halved, so the iterates differ from Boost's but the result carries the same
// guarantee: the midpoint of a bracket of relative width 2^-10 around the root.  NaN where the
// reference's solver throws (the price outside the no-arbitrage range, or no bracket within the
// evaluations).
__device__ inline double implied_vol_reference_f64(double price, double K, double tau) {
    bool call = true;
    double target = price;
    if (1.0 > K) {
        call = false;
        target = price - 1.0 + K;
    }
    const double upper = call ? 1.0 : K;
    if (!(target > 0.0) || !(target < upper) || !(tau > 0.0)) return __builtin_nan("");
    double a = 0.2, b = 0.2;
    double fa = target - bs_price_f64(K, tau, a, call), fb = fa;
    int count = 20 - 1;
    if (fa < 0.0) {
        // the root lies below the guess: halve
        while ((fa < 0.0) == (fb < 0.0)) {
            if (count-- == 0) return __builtin_nan("");
            b = a;
            fb = fa;
            a *= 0.5;
            fa = target - bs_price_f64(K, tau, a, call);
        }
    } else {
        while ((fa < 0.0) == (fb < 0.0)) {
            if (count-- == 0) return __builtin_nan("");
            a = b;
            fa = fb;
            b *= 2.0;
            fb = target - bs_price_f64(K, tau, b, call);
        }
    }
    // f(a) >= 0 > f(b) (or a root on an end point): narrow [a, b] to the tolerance
    for (int it = 0; it < 200 && !(fabs(b - a) <= 0x1p-10 * fmin(fabs(a), fabs(b))); ++it) {
        if (fa == 0.0) return a;
        if (fb == 0.0) return b;
        const double m = a + 0.5 * (b - a);
        const double fm = target - bs_price_f64(K, tau, m, call);
        if ((fm < 0.0) == (fa < 0.0)) {
            a = m;
            fa = fm;
        } else {
            b = m;
            fb = fm;
        }
    }
    return a + (b - a) / 2;
}

}  // namespace rb
