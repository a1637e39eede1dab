// rb_turbo.hip -- the turbo-charged Romano-Touzi estimator (FP64): a control variate and
// antithetic pairs, following ComputePriceRTVarRed (ML_console_ST/RBergomi.cpp:219-288),
// estimateControlVariate (:493-537) and updatePayoffControlVariate (:478-491).
//
// Per option i (Iv, Is the integrals of its group, S' = exp(-rho^2 Iv / 2 + rho Is)):
//   X = C_BS(S', K, 1, sqrt((1 - rho^2) Iv))      the Romano-Touzi payoff
//   Y = C_BS(S', K, 1, sqrt(rho^2 (Q - Iv)))      the control variate, E[Y] = C_BS(1, K, 1, sqrt(rho^2 Q))
// Q = max over the pilot paths of 2 Iv (k_turbo_q), alpha = -cov(X, Y) / cov(Y, Y) over the pilot
// paths with two-pass sample covariances (k_turbo_alpha; ML_console_ST/aux.h:207-220), the main
// payoff (X + alpha Y)/2 on W and on -W (k_turbo_payoff; qmax replaces Q on paths with Q <= Iv,
// which are counted), and price = mean - alpha E[Y] with the reference's exception weighting
// (k_turbo_adjust).  Every sum runs in a fixed order (deterministic results).
#include <hip/hip_runtime.h>

#include "rb_device.h"
#include "rb_kernels.h"
#include "rb_types.h"

namespace rb {

// C_BS(S, K, tau = 1, sigma), S given with its logarithm
__device__ __forceinline__ double bs_call_ln(double lnS, double S, double K, double lnK, double sig) {
    const double d1 = (lnS - lnK + 0.5 * sig * sig) / sig;
    return phi_f64(d1) * S - phi_f64(d1 - sig) * K;
}

__device__ __forceinline__ double block_sum_256(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Q[g] = max over the pilot paths of 2 Iv.  Grid: groups; block 256.
__global__ void __launch_bounds__(256) k_turbo_q(const double* __restrict__ ints, long np, double* __restrict__ Qg) {
    __shared__ double sh[256];
    const int g = blockIdx.x;
    double m = 0.0;
    for (long p = threadIdx.x; p < np; p += 256) m = fmax(m, 2.0 * ints[(long(g) * np + p) * 2]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) Qg[g] = sh[0];
}

__device__ __forceinline__ void xy_of(const double* ints, long np, int g, long p, const Option& o,
                                      double Q, double& X, double& Y) {
    const double Iv = ints[(long(g) * np + p) * 2], Is = ints[(long(g) * np + p) * 2 + 1];
    const double r2 = o.rho * o.rho;
    const double lnS = -0.5 * r2 * Iv + o.rho * Is;
    const double S = exp(lnS);
    X = bs_call_ln(lnS, S, o.K, o.lnK, sqrt((1.0 - r2) * Iv));
    Y = bs_call_ln(lnS, S, o.K, o.lnK, sqrt(r2 * (Q - Iv)));
}

// alpha[i] = -cov(X, Y) / cov(Y, Y) over the pilot paths (0 when cov(Y, Y) = 0).  Grid: options.
__global__ void __launch_bounds__(256)
k_turbo_alpha(const double* __restrict__ ints, long np, const Option* __restrict__ opts,
              const double* __restrict__ Qg, double* __restrict__ alpha) {
    __shared__ double sh[256];
    const int I = blockIdx.x;
    const Option o = opts[I];
    const double Q = Qg[o.group];
    double sx = 0.0, sy = 0.0;
    for (long p = threadIdx.x; p < np; p += 256) {
        double X, Y;
        xy_of(ints, np, o.group, p, o, Q, X, Y);
        sx += X;
        sy += Y;
    }
    const double mx = block_sum_256(sx, sh) / double(np);
    const double my = block_sum_256(sy, sh) / double(np);
    double cxy = 0.0, cyy = 0.0;
    for (long p = threadIdx.x; p < np; p += 256) {
        double X, Y;
        xy_of(ints, np, o.group, p, o, Q, X, Y);
        cxy += (X - mx) * (Y - my);
        cyy += (Y - my) * (Y - my);
    }
    const double vxy = block_sum_256(cxy, sh) / double(np - 1);
    const double vyy = block_sum_256(cyy, sh) / double(np - 1);
    if (threadIdx.x == 0) alpha[I] = vyy > 0.0 ? -vxy / vyy : 0.0;
}

// One evaluation of X + alpha Y; exc = Q <= Iv (qmax used instead)
__device__ __forceinline__ double cv_payoff(double Iv, double Is, const Option& o, double Q,
                                            double a, double qmax, int& exc) {
    const double r2 = o.rho * o.rho;
    const double lnS = -0.5 * r2 * Iv + o.rho * Is;
    const double S = exp(lnS);
    const double X = bs_call_ln(lnS, S, o.K, o.lnK, sqrt((1.0 - r2) * Iv));
    const bool e = !(Q > Iv);
    exc += e ? 1 : 0;
    const double Y = bs_call_ln(lnS, S, o.K, o.lnK, sqrt(r2 * ((e ? qmax : Q) - Iv)));
    return X + a * Y;
}

// Main paths: pair payoff (f(W) + f(-W))/2 per (path, option), block sums of it and its square
// (fixed tree) and of the exception count.  Grid: (path blocks of 256, options).
__global__ void __launch_bounds__(256)
k_turbo_payoff(const double* __restrict__ ints_p, const double* __restrict__ ints_m, long np,
               const Option* __restrict__ opts, const double* __restrict__ Qg,
               const double* __restrict__ alpha, double qmax, long blk0, long nblk_total,
               double* __restrict__ partial, int* __restrict__ exc) {
    __shared__ double s1[256], s2[256];
    __shared__ int se[256];
    const int I = blockIdx.y;
    const Option o = opts[I];
    const long p = long(blockIdx.x) * 256 + threadIdx.x;
    double x = 0.0;
    int e = 0;
    if (p < np) {
        const double Q = Qg[o.group], a = alpha[I];
        const long k = (long(o.group) * np + p) * 2;
        x = 0.5 * cv_payoff(ints_p[k], ints_p[k + 1], o, Q, a, qmax, e);
        x += 0.5 * cv_payoff(ints_m[k], ints_m[k + 1], o, Q, a, qmax, e);
    }
    s1[threadIdx.x] = x;
    s2[threadIdx.x] = x * x;
    se[threadIdx.x] = e;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
            se[threadIdx.x] += se[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* d = partial + (long(I) * nblk_total + blk0 + blockIdx.x) * 2;
        d[0] = s1[0];
        d[1] = s2[0];
        exc[long(I) * nblk_total + blk0 + blockIdx.x] = se[0];
    }
}

// price -= alpha [(1 - p) C_BS(1, K, 1, sqrt(rho^2 Q)) + p C_BS(1, K, 1, sqrt(rho^2 qmax))],
// p = exceptions / M (ML_console_ST/RBergomi.cpp:276-283).  One thread per option.
__global__ void k_turbo_adjust(double* __restrict__ price, const Option* __restrict__ opts, int nopt,
                               const double* __restrict__ Qg, const double* __restrict__ alpha,
                               double qmax, const int* __restrict__ exc, long nblk, long M,
                               long* __restrict__ count) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nopt) return;
    const Option o = opts[I];
    long c = 0;
    for (long b = 0; b < nblk; ++b) c += exc[long(I) * nblk + b];
    const double pq = double(c) / double(M);
    const double r2 = o.rho * o.rho;
    const double adj = (1.0 - pq) * bs_call_ln(0.0, 1.0, o.K, o.lnK, sqrt(r2 * Qg[o.group])) +
                       pq * bs_call_ln(0.0, 1.0, o.K, o.lnK, sqrt(r2 * qmax));
    price[I] -= alpha[I] * adj;
    count[I] = c;
}

hipError_t launch_turbo_q(const double* ints, long np, int ngroups, double* Qg, hipStream_t st) {
    hipLaunchKernelGGL(k_turbo_q, dim3(ngroups), dim3(256), 0, st, ints, np, Qg);
    return hipGetLastError();
}

hipError_t launch_turbo_alpha(const double* ints, long np, const Option* opts, int nopt,
                              const double* Qg, double* alpha, hipStream_t st) {
    hipLaunchKernelGGL(k_turbo_alpha, dim3(nopt), dim3(256), 0, st, ints, np, opts, Qg, alpha);
    return hipGetLastError();
}

hipError_t launch_turbo_payoff(const double* ints_p, const double* ints_m, long np,
                               const Option* opts, int nopt, const double* Qg,
                               const double* alpha, double qmax, long blk0, long nblk_total,
                               double* partial, int* exc, hipStream_t st) {
    dim3 grid(unsigned((np + 255) / 256), unsigned(nopt));
    hipLaunchKernelGGL(k_turbo_payoff, grid, dim3(256), 0, st, ints_p, ints_m, np, opts, Qg, alpha,
                       qmax, blk0, nblk_total, partial, exc);
    return hipGetLastError();
}

hipError_t launch_turbo_adjust(double* price, const Option* opts, int nopt, const double* Qg,
                               const double* alpha, double qmax, const int* exc, long nblk,
                               long M, long* count, hipStream_t st) {
    hipLaunchKernelGGL(k_turbo_adjust, dim3((nopt + 63) / 64), dim3(64), 0, st, price, opts, nopt,
                       Qg, alpha, qmax, exc, nblk, M, count);
    return hipGetLastError();
}

}  // namespace rb
