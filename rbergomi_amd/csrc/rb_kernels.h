// rb_kernels.h -- host-side launchers of the HIP kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "rb_device.h"
#include "rb_types.h"

namespace rb {

// raise a kernel's dynamic-LDS limit (once per device and size)
hipError_t allow_lds(const void* fn, size_t bytes);

// staged FP64 path (rb_staged.hip)
hipError_t launch_normals_f64(PhiloxKey key, long p0, long np, int nstreams, int N, double* out,
                              hipStream_t st);
size_t volterra_f64_lds_bytes(int N);
hipError_t launch_volterra_f64(const double* Z, long zstride, long np, int N, int nunits,
                               const Unit* units, const double* gamma, double* Wt,
                               hipStream_t st);
// the same convolution on the FP64 matrix cores, N <= 1024 (rb_volterra64.hip)
constexpr int VOLTERRA_MFMA_MAX_N = 1024;
size_t volterra_mfma_lds_bytes(int N);
hipError_t launch_volterra_mfma_f64(const double* Z, long zstride, long np, int N, int nunits,
                                    const Unit* units, const double* gamma, double* Wt, hipStream_t st);
// beyond: the W1 tile streamed through LDS in segments (segment length for an LDS limit)
int volterra_mfma_long_segment(int N, size_t lds_limit);
size_t volterra_mfma_long_lds_bytes(int N, int S);
hipError_t launch_volterra_mfma_long_f64(const double* Z, long zstride, long np, int N, int nunits,
                                         const Unit* units, const double* gamma, double* Wt,
                                         size_t lds_limit, hipStream_t st);
// FP64 convolution of the staged path: on the matrix cores (whole W1 tile in LDS up to
// VOLTERRA_MFMA_MAX_N, segmented beyond while a 16-step segment fits), else the direct sum
inline int volterra_conv_kind(int N, size_t lds_limit) {
    return N <= VOLTERRA_MFMA_MAX_N ? 0 : volterra_mfma_long_segment(N, lds_limit) >= 16 ? 1 : 2;
}
inline hipError_t launch_volterra_conv_f64(const double* Z, long zstride, long np, int N, int nunits,
                                           const Unit* units, const double* gamma, double* Wt,
                                           size_t lds_limit, hipStream_t st) {
    switch (volterra_conv_kind(N, lds_limit)) {
        case 0: return launch_volterra_mfma_f64(Z, zstride, np, N, nunits, units, gamma, Wt, st);
        case 1: return launch_volterra_mfma_long_f64(Z, zstride, np, N, nunits, units, gamma, Wt, lds_limit, st);
        default: return launch_volterra_f64(Z, zstride, np, N, nunits, units, gamma, Wt, st);
    }
}
hipError_t launch_scan_f64(int method, const double* Z, long zstride, const double* Wt, long np,
                           int N, const Group* groups, int ngroups, const double* comp,
                           double* out, hipStream_t st, double sign = 1.0);
hipError_t launch_payoff_f64(const double* ints, long np, long p0, const Option* opts,
                             int out_size, int method, long blk0, long nblk_total,
                             double* partial, double* payoff, hipStream_t st);
hipError_t launch_reduce_f64(const double* partial, long nblk, long M, int out_size,
                             double* price, double* stat, hipStream_t st);
hipError_t launch_reduce_rows(const double* partial, const long* row_off, const long* row_n,
                              const long* row_M, int nrows, double* price, double* stat,
                              hipStream_t st);
// launch_reduce_rows and launch_iv_f64 in one kernel, which also sets *counter to zero (nrows >= 1)
hipError_t launch_reduce_rows_iv(const double* partial, const long* row_off, const long* row_n,
                                 const long* row_M, const Option* opts, int nrows, double* price,
                                 double* stat, double* iv, int* counter, hipStream_t st);
hipError_t launch_iv_ref_f64(const double* price, const double* K, const double* T, int n, double* iv,
                             hipStream_t st);
hipError_t launch_iv_f64(const double* price, const Option* opts, int n, double* iv,
                         hipStream_t st);
hipError_t launch_philox_probe(const void* ctr, PhiloxKey key, long n, void* out, hipStream_t st);
// MFMA layout probe (rb_probe.hip)
hipError_t launch_probe_mfma_layout(const float* A, const float* B, float* out, int mode,
                                    hipStream_t st);

hipError_t launch_probe_payoff(const double* lnS, const double* sig, const double* K, int n,
                               double* out, float* mills_out, const float* z, int form, hipStream_t st);

hipError_t launch_probe_copy(const void* src, void* dst, long bytes, hipStream_t st);
hipError_t launch_probe_issue(int op, float* out, uint32_t seed, hipStream_t st);
double probe_issue_insts_per_simd(int simds);

// FFT convolution for long paths (rb_fft.hip)
size_t volterra_fft_lds_bytes(int log2L);
hipError_t launch_volterra_fft(const double* Z, long zstride, long np, int N, int log2L, int nunits,
                               const Unit* units, const float2* ghat, const float2* tw, double* Wt,
                               hipStream_t st);

// Sobol' Gaussians (quasi-Monte Carlo), same [path][stream][N] layout as the Philox ones
hipError_t launch_normals_sobol_f64(const uint32_t* V, const uint32_t* shift, long p0, long np,
                                    int nstreams, int N, double* out, hipStream_t st);
hipError_t launch_norm_inv_probe(const double* u, long n, double* z, hipStream_t st);
// Brownian-bridge (hierarchical) mapping of rows of N = 2^m Gaussians, in place
hipError_t launch_bridge_f64(double* Z, long rows, int N, hipStream_t st);
// exact scheme: Wtilde = [L21 | L22] (x1, x2) per unit
hipError_t launch_volterra_chol(const double* Z, long zstride, long np, int n, int nunits,
                                const double* Lc, double* Wt, hipStream_t st);
// turbo-charged RT estimator (control variate + antithetics)
hipError_t launch_turbo_q(const double* ints, long np, int ngroups, double* Qg, hipStream_t st);
hipError_t launch_turbo_alpha(const double* ints, long np, const Option* opts, int nopt,
                              const double* Qg, double* alpha, hipStream_t st);
hipError_t launch_turbo_payoff(const double* ints_p, const double* ints_m, long np,
                               const Option* opts, int nopt, const double* Qg,
                               const double* alpha, double qmax, long blk0, long nblk_total,
                               double* partial, int* exc, hipStream_t st);
hipError_t launch_turbo_adjust(double* price, const Option* opts, int nopt, const double* Qg,
                               const double* alpha, double qmax, const int* exc, long nblk,
                               long M, long* count, hipStream_t st);

}  // namespace rb
