// rb_engine.h -- host planner and executor of the rough-Bergomi pricing path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "rb_device.h"
#include "rb_fused.h"
#include "rb_fused_host.h"
#include "rb_types.h"

namespace rb {

enum Precision { PREC_FP64 = 0, PREC_FP32 = 1, PREC_SPLIT_FP16 = 2, PREC_SPLIT_BF16 = 3, PREC_FP32_GEMM = 4,
                 PREC_FP32_FFT = 5 };

// operand mode of the fused kernel for a precision (0 FP16 hi+lo, 1 BF16 hi+lo, 2 FP32), -1 if
// the precision does not run on the fused kernel
inline int fused_mode_of(int precision) {
    return precision == PREC_SPLIT_FP16 ? 0 : precision == PREC_SPLIT_BF16 ? 1 : precision == PREC_FP32_GEMM ? 2 : -1;
}
enum Compensator { COMP_REF = 0, COMP_EXACT = 1 };
enum Status {
    ST_OK = 0,
    ST_ERR_ARG = 1,
    ST_ERR_SIZE = 2,
    ST_ERR_DEVICE = 3,
    ST_ERR_NOMEM = 4,
    ST_ERR_LIMIT = 5,
};

// One pricing request in the reference's tensor form: (H, eta, rho) zipped over par_size,
// (T, K) zipped over op_size, outputs I = h*op_size + t (rBergLib/ParamTot.h:32-33).
struct Request {
    double xi = 0.0;
    std::vector<double> H, eta, rho, T, K;
    // unordered rows: output i prices (T[i], K[i]) under (xi_rows[i], H[i], eta[i], rho[i]);
    // otherwise the reference's zipped tensor form, output h*|T| + t
    bool rows = false;
    std::vector<double> xi_rows;
    int N = 0;
    long M = 0;
    std::vector<uint64_t> seed;
    int method = METHOD_RT;
    int precision = PREC_FP64;
    int comp = COMP_REF;
    // turbo-charged RT estimator (control variate + antithetics, ML_console_ST/RBergomi.cpp:219-288):
    // m_control pilot paths estimate Q and alpha; q_max replaces Q where Q <= Iv
    int scheme = 0;              // 0 hybrid scheme, 1 exact Cholesky sampling of Wtilde
    int gauss = 0;               // Gaussian source: 0 Philox (seeded), 1 Sobol' points (QMC),
                                 // 2 Sobol' points through the Brownian-bridge mapping
    bool turbo = false;
    long m_control = 10000;
    double q_max = 10000.0;
    int out_size() const { return rows ? int(H.size()) : int(H.size() * T.size()); }
};

// Deduplicated work tables (see rb_types.h)
struct Layout {
    int N = 0;
    std::vector<Unit> units;
    std::vector<double> gamma;  // [units][N]
    std::vector<Group> groups;
    std::vector<double> comp;   // [groups][N]: c_i = eta^2/2 ((i-1) dt)^(2H), c_0 = 0
    std::vector<Option> options;
};

enum Scheme { SCHEME_HYBRID = 0, SCHEME_CHOLESKY = 1 };
// exact sampling: Gauss hypergeometric 2F1 for 0 <= z <= 1 (c - a - b non-integer), the joint
// covariance blocks A21 = Cov(Wtilde, W1), A22 = Cov(Wtilde, Wtilde) (n x n each), and the factor
// L = [A21 | chol(A22 - A21 A21^T)] (n rows of 2n)
double hyp2f1(double a, double b, double c, double z);
void chol_covariance(double H, int n, double* A21, double* A22);
int chol_factor(double H, int n, std::vector<double>& L, std::string& err);

// Sobol' direction integers V[d*32 + k] of dimensions d < dims (rb_sobol.cpp)
void sobol_directions(int dims, std::vector<uint32_t>& V);
std::vector<uint64_t> primitive_polynomials(int count);

// Philox path counters of the turbo-charged estimator's pilot paths (disjoint from main paths)
constexpr long TURBO_PILOT_BASE = 1L << 62;

int validate(const Request& r, std::string& err);
Layout make_layout(const Request& r);
PhiloxKey philox_key(const uint64_t* seed, int n);
// FFT in FP64 on the host (radix-2, a.size() a power of two; sign -1 forward, unnormalised)
void fft_host(std::vector<std::complex<double>>& a, int sign);
// i with its low `bits` bits in reverse order
inline int bit_reverse(int i, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

// Per-kernel timing accumulated over the launches of one run (HIP events on the run's stream).
struct KernelTime {
    std::string name;
    double total_ms = 0.0;
    int launches = 0;
};

// Device-resident plan: tables uploaded once, chunk buffers allocated once, run many times.
struct Plan {
    Request req;
    Layout lay;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool profile = false;
    long chunk = 0;          // paths per chunk
    long nblk = 0;           // 256-path blocks over all M
    int nstreams = 2;        // Gaussian blocks per path: W1, W1perp (+ Wperp for plain MC)
    // device tables
    Unit* d_units = nullptr;
    double* d_gamma = nullptr;
    Group* d_groups = nullptr;
    double* d_comp = nullptr;
    Option* d_opts = nullptr;
    // device work buffers
    double* d_Z = nullptr;        // [chunk][nstreams][N]
    double* d_Wt = nullptr;       // [units][chunk][N]
    double* d_ints = nullptr;     // [groups][chunk][2]
    double* d_partial = nullptr;  // [options][nblk][2]
    double* d_price = nullptr;
    double* d_stat = nullptr;
    double* d_iv = nullptr;
    double* d_payoff = nullptr;   // [chunk][options], quadrature mode only
    // turbo-charged estimator
    double* d_ints2 = nullptr;    // [groups][chunk][2] integrals of the antithetic paths
    double* d_Qg = nullptr;       // [groups] Q = max over the pilot paths of 2 Iv
    double* d_chol = nullptr;     // [units][N][2N] Cholesky factors (exact scheme)
    uint32_t* d_sobol = nullptr;  // [streams*N][32] Sobol' direction integers (QMC)
    uint32_t* d_shift = nullptr;  // [streams*N] digital shift (seeded QMC), or null
    double* d_alpha = nullptr;    // [options] control-variate coefficients
    int* d_exc = nullptr;         // [options][nblk] Q <= Iv evaluations per path block
    long* d_count = nullptr;      // [options] their totals
    size_t lds_limit = 65536;     // LDS bytes one workgroup may use on the plan's device
    // long-path FP32 FFT convolution (replaces the FP64 convolution in the staged pipeline)
    bool conv_fft = false;
    int log2L = 0;
    float2* d_tw = nullptr;       // exp(-2 pi i k / L), k < L/2
    float2* d_ghat = nullptr;     // [units][L] FFT of the zero-padded Gamma
    // fused mixed-precision path (a batch of one request)
    bool fused = false;
    int fused_mode = 0;           // 0 = FP16 split, 1 = BF16 split
    std::unique_ptr<FusedHostReq> fused_req;
    std::unique_ptr<FusedBatch> fused_batch;
    // long-path mixed-precision path (single-launch FFT kernel, Romano-Touzi)
    std::unique_ptr<FftFused> fused_fft;
    // graph replay: the first run is captured into a hipGraph, later runs relaunch it
    bool use_graph = false;
    bool capturing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<KernelTime> times;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;

    ~Plan();
    int init(const Request& r, hipStream_t st, bool quadrature, std::string& err);
    // seeded run: Philox Gaussians on the device
    int run(std::string& err);
    // user Gaussians (host buffer, quadrature layout), per-path payoffs back to the host
    int run_samples(const double* Z_host, double* payoff_host, std::string& err);
    int results(double* price, double* iv, double* stat, std::string& err);
    // turbo-charged estimator: per-option alpha, Q and count of Q <= Iv evaluations (each may be null)
    int turbo_results(double* alpha, double* Q, long* exceptions, std::string& err);
    int init_fused(std::string& err);
    int init_fft(std::string& err);
    hipError_t conv(long np, long zstride);
    // enqueue one run on the plan's stream, no synchronization
    int enqueue(std::string& err);
    void begin_timing();
    void end_timing();
    int timed(int id);
    void mark(int id, hipEvent_t a, hipEvent_t b);
};

// Kernel ids for timing
enum KernelId { K_NORMALS = 0, K_VOLTERRA, K_SCAN, K_PAYOFF, K_REDUCE, K_IV, K_FUSED, K_COUNT };
const char* kernel_name(int id);

// Device properties, queried once per device (the query is slow next to a small request).
hipError_t device_props(int device, hipDeviceProp_t& out);
// Device buffers of plans, reused across plans of the process (see rb_engine.cpp)
hipError_t ws_alloc(void** p, size_t bytes);
void ws_free(void* p);
// Keeps up to 64 GiB of freed stream-ordered memory in the device's default pool (set once per
// device), so that plans and batches reuse it instead of remapping.
void keep_pool_cached(int device);

// thread-local error channel of the C API
void set_last_error(const std::string& s);
const char* last_error();

}  // namespace rb
