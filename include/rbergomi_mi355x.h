/*
 * rbergomi_mi355x.h -- C ABI of librbergomi_mi355x.so, the MI355X (gfx950) rough-Bergomi pricer.
 *
 * Part 1 is the drop-in boundary: the three extern "C" entry points of the reference's
 * rBergLib C interface with identical signatures, argument meaning and error behaviour.
 * Part 2 adds status-returning entry points (v2), device-resident plans and probes used by
 * the parity tests and the benchmark.
 *
 * Conventions shared with the reference (rBergLib/cinterface.h:17-60):
 *   (H, eta, rho) are zipped arrays of length par_size, (T, K) zipped arrays of length
 *   op_size; every model triple is priced for every option pair; output index
 *   I = h*op_size + t, out_size = par_size*op_size; S0 = 1, r = 0, constant forward variance xi.
 */
#ifndef RBERGOMI_MI355X_H
#define RBERGOMI_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================================
 * Part 1 -- reference C ABI (drop-in)
 * ====================================================================================== */

/* Plain Monte Carlo prices and implied vols (log-Euler spot, payoff (S_T-K)^+).
 * Replaces computeIV_tensor, rBergLib/cinterface.h:63-66 (implementation
 * rBergLib/cinterface.cpp:10-46).  Like the reference it writes the rho values into the
 * eta_tot buffer and the eta values into rho_tot (the swapped Result2Array call at
 * rBergLib/cinterface.cpp:45); rb_price_v2 does not swap.
 * out_size != par_size*op_size: message on stderr and exit(17); num_threads < 1: exit(21).
 * num_threads >= 1 is accepted as a hint only: the results do not depend on it.
 * A device failure prints a message and exits with status 40 (the reference ABI has no
 * error channel). */
void computeIV_tensor(double xi, double* H, double* eta, double* rho, int par_size,
        double* T, double* K, int op_size, int N, long M, uint64_t* seed, int seed_size,
        int num_threads, double* price, double* iv, double* stat, double* H_tot, double* eta_tot,
        double* rho_tot, double* T_tot, double* K_tot, int out_size);

/* Romano-Touzi conditioned prices and implied vols: the hot path.
 * Replaces computeIVRT_tensor, rBergLib/cinterface.h:68-71 (rBergLib/cinterface.cpp:49-85).
 * Same error behaviour as computeIV_tensor; the echo arrays are not swapped. */
void computeIVRT_tensor(double xi, double* H, double* eta, double* rho, int par_size,
        double* T, double* K, int op_size, int N, long M, uint64_t* seed, int seed_size,
        int num_threads, double* price, double* iv, double* stat, double* H_tot, double* eta_tot,
        double* rho_tot, double* T_tot, double* K_tot, int out_size);

/* Romano-Touzi per-path payoffs from user Gaussians.
 * Replaces computeIVRT_tensor_quadrature, rBergLib/cinterface.h:73-75
 * (rBergLib/cinterface.cpp:102-131).  Z holds 2*N*M doubles: path m uses
 * W1 = Z[2mN, 2mN+N) and W1perp = Z[2mN+N, 2mN+2N).  payoff holds M*out_size doubles,
 * payoff[m*out_size + I].  No size check (as in the reference). */
void computeIVRT_tensor_quadrature(double xi, double* H, double* eta, double* rho, int par_size,
        double* T, double* K, int op_size, int N, long M, double* Z,
        int num_threads, double* payoff);

/* ======================================================================================
 * Part 2 -- status-returning API (no counterpart in the reference)
 * ====================================================================================== */

enum rb_status {
    RB_OK = 0,
    RB_ERR_ARG = 1,     /* invalid argument values                    */
    RB_ERR_SIZE = 2,    /* out_size != par_size*op_size               */
    RB_ERR_DEVICE = 3,  /* HIP runtime or kernel failure              */
    RB_ERR_NOMEM = 4,   /* device allocation failed                   */
    RB_ERR_LIMIT = 5    /* size beyond what this build supports       */
};
enum rb_precision {
    RB_PREC_FP64 = 0,        /* every stage in FP64 (the parity path)               */
    RB_PREC_FP32 = 1,        /* Volterra convolution by FP32 FFT in LDS, the rest FP64 */
    RB_PREC_SPLIT_FP16 = 2,  /* Toeplitz GEMM as FP16 hi+lo products, FP32 accumulate */
    RB_PREC_SPLIT_BF16 = 3,  /* Toeplitz GEMM as BF16 hi+lo products, FP32 accumulate */
    RB_PREC_FP32_GEMM = 4,   /* Toeplitz GEMM on FP32 operands (v_mfma_f32_32x32x2_f32), N <= 1024 */
    RB_PREC_FP32_FFT = 5     /* FP32 LDS FFT convolution inside the single-launch kernel at every
                                N <= 8192 (the long-path kernel), FP32 integrals, FP64 sums;
                                caller or Sobol' Gaussians take the staged FFT pipeline; not
                                accepted by rb_price_batch and the server */
};
enum rb_compensator {
    RB_COMP_REF = 0,   /* 0.5 eta^2 ((i-1) dt)^(2H), rBergLib/rBergomiMT.cpp:77-79  */
    RB_COMP_EXACT = 1  /* 0.5 eta^2 (i dt)^(2H), clean_Rfiles/hybridScheme.R:110-113 */
};
enum rb_method { RB_METHOD_RT = 0, RB_METHOD_MC = 1 };
/* How Wtilde is sampled.  HYBRID: the reference's hybrid scheme (kappa = 1).  CHOLESKY: exactly,
 * by the Cholesky factor of the joint covariance of (W1, Wtilde) (RfBm, PremiaRBergomi/RfBm.cpp:
 * 18-95), FP64 path only, H in (0, 1/2), N <= 2048; the path's two Gaussian blocks are the
 * factor's inputs (the first is W1).  It has no discretisation bias in Wtilde, so it measures the
 * hybrid scheme's bias against N. */
enum rb_scheme { RB_SCHEME_HYBRID = 0, RB_SCHEME_CHOLESKY = 1 };

typedef struct rb_config {
    int precision;    /* rb_precision   */
    int compensator;  /* rb_compensator */
    int method;       /* rb_method      */
    int scheme;       /* rb_scheme (0 in older callers' zero-initialised configs: hybrid) */
} rb_config;

/* Message of the last failed call on this thread ("" if none). */
const char* rb_last_error(void);

/* Prices, implied vols and standard errors; returns an rb_status.  cfg may be NULL
 * (FP64, reference compensator, Romano-Touzi).  iv may be NULL. */
int rb_price_v2(double xi, const double* H, const double* eta, const double* rho, int par_size,
                const double* T, const double* K, int op_size, int N, long M,
                const uint64_t* seed, int seed_size, const rb_config* cfg,
                double* price, double* iv, double* stat, int out_size);

/* Per-path payoffs from user Gaussians (quadrature layout of computeIVRT_tensor_quadrature:
 * 2 blocks of N per path for RB_METHOD_RT, 3 blocks W1, W1perp, Wperp for RB_METHOD_MC).
 * price/stat (may be NULL) receive the sample mean and standard error. */
int rb_payoff_samples_v2(double xi, const double* H, const double* eta, const double* rho,
                         int par_size, const double* T, const double* K, int op_size, int N,
                         long M, const double* Z, const rb_config* cfg, double* payoff,
                         double* price, double* stat);

/* Device-resident plan: validated, tables uploaded and work buffers allocated once;
 * rb_plan_run launches the whole pricing pipeline on the plan's stream (stream = NULL
 * creates one; otherwise a hipStream_t the caller owns) and waits for it. */
typedef struct rb_plan rb_plan;
int rb_plan_create(double xi, const double* H, const double* eta, const double* rho,
                   int par_size, const double* T, const double* K, int op_size, int N, long M,
                   const uint64_t* seed, int seed_size, const rb_config* cfg, void* stream,
                   rb_plan** out);
int rb_plan_run(rb_plan* plan);
int rb_plan_results(rb_plan* plan, double* price, double* iv, double* stat);
/* Graph replay (enable = 1): the next rb_plan_run records the plan's launches into a hipGraph
 * and every run replays it (no per-launch host overhead; kernel timing is off in this mode). */
int rb_plan_set_graph(rb_plan* plan, int enable);
/* Kernel timing of the following runs with HIP events on the plan's stream (enable = 1). */
int rb_plan_set_profiling(rb_plan* plan, int enable);
/* Per-kernel totals of the last run: names (NUL-terminated, 32 bytes each), total ms and
 * launch counts for up to max entries; *count receives the number written. */
int rb_plan_kernel_times(rb_plan* plan, char* names, double* total_ms, int* launches, int max,
                         int* count);
/* Work sizes of the plan: units (distinct H), groups, options, paths per chunk. */
int rb_plan_info(rb_plan* plan, int* nunits, int* ngroups, int* noptions, long* chunk);
void rb_plan_destroy(rb_plan* plan);

/* Quasi-Monte Carlo: path m (0-based) uses Sobol' point m + 1 in 2N dimensions (3N for plain
 * MC), mapped by the inverse normal CDF; dimension 2j goes to W1_j and 2j+1 to W1perp_j (breakZ),
 * as ComputePriceRTMT_sobol / ComputeIVRTMT_sobol (rBergomi/rBergomiMT.cpp:610-716) with
 * normalQMC_sample (rBergomi/qmc.cpp:26-41).  The direction numbers are this library's own
 * (primitive polynomials by degree, seeded odd initial numbers; the reference's table is not
 * reproduced), so the points differ from the reference's beyond the first dimension.  With
 * seed_size = 0 this is the reference's construction (points 1..M, unscrambled; such an
 * incomplete net carries a bias of the order of the Monte Carlo error at moderate M).  With a
 * seed, a random digital shift keyed by the seed is XORed into every dimension and points 0..M-1
 * are used at cell midpoints: an unbiased randomised QMC estimate whose error is estimated from
 * independent seeds.  M <= 2^32 - 2.  cfg as for rb_price_v2 (the fused-kernel precisions
 * take the FP32-FFT staged path). */
int rb_price_sobol(double xi, const double* H, const double* eta, const double* rho, int par_size,
                   const double* T, const double* K, int op_size, int N, long M,
                   const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                   double* iv, double* stat, int out_size);
/* The same points through the Brownian-bridge (hierarchical) mapping of each stream's N = 2^m
 * dimensions, hierarchical2increments (rBergomi/aux.h:105-125; used with Sobol' points in
 * rBergomi/text/tests.cpp:139-140): the first dimensions carry the coarse shape of each path. */
int rb_price_sobol_bridge(double xi, const double* H, const double* eta, const double* rho, int par_size,
                          const double* T, const double* K, int op_size, int N, long M,
                          const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                          double* iv, double* stat, int out_size);

/* ---- unordered rows ----
 * Row i prices the call (T[i], K[i]) under (xi[i], H[i], eta[i], rho[i]); every row uses the same
 * Gaussians of each path, and outputs are indexed by row.  Replaces ComputePriceRTMTunstructured /
 * ComputeIVRTMTunstructured (rBergomi/rBergomiMT.cpp:837-935) over ParamTotUnordered
 * (rBergomi/ParamTot.h:57-75), whose xi is one scalar; the per-row xi is that of the ParamTot of
 * ML_console_ST/ParamTot.h:25-43 (pass the same value in every row for the former).  cfg as for
 * rb_price_v2 (NULL: FP64, Romano-Touzi, reference compensator); at most 65535 rows. */
int rb_price_rows(const double* xi, const double* H, const double* eta, const double* rho,
                  const double* T, const double* K, int rows, int N, long M, const uint64_t* seed,
                  int seed_size, const rb_config* cfg, double* price, double* iv, double* stat);
int rb_plan_create_rows(const double* xi, const double* H, const double* eta, const double* rho,
                        const double* T, const double* K, int rows, int N, long M,
                        const uint64_t* seed, int seed_size, const rb_config* cfg, void* stream,
                        rb_plan** out);
/* Turbo-charged Romano-Touzi estimator of unordered rows, FP64: a control variate and antithetic
 * pairs, replacing ComputePriceRTVarRed (ML_console_ST/RBergomi.cpp:219-288, with
 * estimateControlVariate :493-537 and updatePayoffControlVariate :478-491).  m_control pilot paths
 * (the reference's Mcontrol, 10000 there; drawn from Philox counters disjoint from the M main
 * paths) give Q_i = max 2 Iv and alpha_i = -cov(X, Y)/cov(Y, Y); each main path prices
 * (X + alpha Y)/2 on (W1, W1perp) and on (-W1, -W1perp), with q_max (the reference's Qmax, 10000)
 * in place of Q where Q <= Iv; price = mean - alpha E[Y].  stat is the standard error of the
 * pair-averaged payoff.  alpha, Q (per row) and exceptions (Q <= Iv evaluations per row) may be
 * NULL.  Requires m_control <= the paths of one chunk (about 6 GB of staged buffers). */
int rb_price_rows_turbo(const double* xi, const double* H, const double* eta, const double* rho,
                        const double* T, const double* K, int rows, int N, long M,
                        const uint64_t* seed, int seed_size, long m_control, double q_max,
                        double* price, double* iv, double* stat, double* alpha, double* Q,
                        long* exceptions);
/* Caller Gaussians in the quadrature layout (path m: W1 = Z[2mN, 2mN+N), W1perp = the next N
 * values; plain MC: 3N values per path, the third block Wperp) -> payoff[m*rows + i], and
 * optionally price/stat (FP64 path). */
int rb_payoff_rows_samples(const double* xi, const double* H, const double* eta, const double* rho,
                           const double* T, const double* K, int rows, int N, long M,
                           const double* Z, const rb_config* cfg, double* payoff, double* price,
                           double* stat);

/* ---- batches and serving (no counterpart in the reference) ----
 * One pricing request in the reference's tensor form plus its caller-owned outputs
 * (out_size = par_size*op_size doubles each; iv may be NULL). */
typedef struct rb_request {
    double xi;
    const double* H;
    const double* eta;
    const double* rho;
    int par_size;
    const double* T;
    const double* K;
    int op_size;
    int N;
    long M;
    const uint64_t* seed;
    int seed_size;
    double* price;
    double* iv;
    double* stat;
} rb_request;

/* Prices n requests with one launch of the fused kernel (cfg->precision must be
 * RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM; requests with 1024 < N <= 8192 take a
 * launch of the long-path FFT kernel each); cfg->method and cfg->compensator apply to
 * every request).  Every request's results are bit-identical to pricing it alone with
 * rb_price_v2 and the same cfg.  The server prices Romano-Touzi requests. */
int rb_price_batch(const rb_request* reqs, int n, const rb_config* cfg);

/* A batch planned once and priced many times, e.g. by a calibration loop that re-prices the same
 * request shapes with new parameters.  No counterpart in the reference, whose driver sets up its
 * OpenMP team and tables on every call (rBergLib/rBergomiMT.cpp:340-355).  rb_batch_create plans
 * the requests, uploads their tables and allocates every buffer; each rb_batch_run prices all of
 * them (short requests in as few fused launches as one workgroup's LDS allows, long ones on the
 * FFT kernel, all on `stream`, NULL: the batch's own) and copies the results into the requests'
 * output arrays, which must stay valid for the batch's life.  Results equal rb_price_batch's. */
typedef struct rb_batch rb_batch;
int rb_batch_create(const rb_request* reqs, int n, const rb_config* cfg, void* stream, rb_batch** out);
/* New parameters (xi, H, eta, rho, T, K, seed) and output arrays for every request, with the same
 * N, M and option counts as at creation and the same table layout (distinct H, maturities and rho
 * per request): the tables of the short requests are copied into the batch's device buffers, so a
 * recorded graph keeps replaying valid launches; long-path requests are rebuilt and the graph is
 * recorded again on the next run.  RB_ERR_ARG if the layout changes. */
int rb_batch_update(rb_batch* b, const rb_request* reqs, int n);
/* Graph replay (enable = 1): the next rb_batch_run records the batch's launches into a hipGraph
 * after one eager run, and every later run replays it (no per-launch host overhead). */
int rb_batch_set_graph(rb_batch* b, int enable);
int rb_batch_run(rb_batch* b);
void rb_batch_destroy(rb_batch* b);

/* Tile-shape autotuning of the fused kernel.  No counterpart in the reference (which has no
 * tiles).  A tile is 32 PB paths by all steps (PB = 1, 2 or 4 path blocks of 32; the four waves
 * of a workgroup split 4 / PB per path block).  For each N[i] (1..1024) this times every valid
 * PB with shape's parameters, options, M and seed (shape->N and its output pointers are ignored)
 * under cfg's precision and method on the current device, and records the fastest for every
 * later plan, batch and server request of the process with that padded N (multiple of 32),
 * operand mode and estimator.  pb_out[i], ms_out[i] (may be NULL): the choice and the time of
 * one launch with it.  Without an entry the default rule applies (PB = 4 for N <= 32 and
 * 65..96, 2 for 33..64 and 97..192, 1 above), measured on MI355X.  Tile shapes change prices
 * only through the order of partial sums: Romano-Touzi prices came out bit-identical for every
 * shape; plain MC also reorders FP32 sums over a wave's row blocks (measured <= 6e-10 in price). */
int rb_autotune_fused(const rb_request* shape, const int* N, int nN, const rb_config* cfg, int* pb_out,
                      double* ms_out);
/* The tuned table as text lines "Np mode mc pb" (mode 0 FP16, 1 BF16, 2 FP32; mc 1 = plain MC).
 * The environment variable RBERGOMI_TUNE_CACHE names such a file, loaded on first use. */
int rb_tune_save(const char* path);
int rb_tune_load(const char* path);
int rb_tune_clear(void);
/* Copies up to max entries as (Np, mode, mc, pb) into rows[4*i..]; returns the entry count. */
int rb_tune_table(int* rows, int max);

/* Continuous-batching server.  A host thread forms batches from the pending requests, admits a
 * request when the device bytes of the batches in flight leave room for it, runs up to
 * max_inflight batches on separate streams (a new batch starts while earlier ones finish), and
 * copies each request's results to its output arrays when its batch completes.  A request with
 * 1024 < N <= 8192 runs as a batch of its own on the long-path FFT kernel. */
typedef struct rb_server rb_server;
typedef struct rb_server_config {
    int precision;            /* RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM */
    int max_inflight;         /* batches on the device at once (>= 1) */
    int max_batch_requests;   /* requests per batch (>= 1) */
    long max_device_bytes;    /* admission budget for the batches in flight */
    double max_batch_path_steps;  /* work cap of one batch, sum of M*N (0: none); a request above
                                   * it runs as a batch of its own.  Short batches bound how long a
                                   * small request waits behind large ones. */
} rb_server_config;
/* Romano-Touzi requests, the reference compensator. */
int rb_server_create(const rb_server_config* cfg, rb_server** out);
/* The same with the estimator (RB_METHOD_RT, or RB_METHOD_MC: computeIV_tensor's plain Monte
 * Carlo) and the compensator every request of the server is priced with. */
int rb_server_create_method(const rb_server_config* cfg, int method, int compensator, rb_server** out);
/* Copies the inputs; the output arrays must stay valid until rb_server_wait returns for the
 * ticket.  Returns RB_ERR_LIMIT for a request that can never fit the device budget. */
int rb_server_submit(rb_server* s, const rb_request* req, long* ticket);
/* Blocks until the request is done; latency_ms (may be NULL) = submit to results copied. */
int rb_server_wait(rb_server* s, long ticket, double* latency_ms);
/* Latency percentiles over the requests completed so far, their count, the paths*steps they
 * priced, and the number of batches launched. */
int rb_server_stats(rb_server* s, double* p50_ms, double* p99_ms, long* completed,
                    double* path_steps, long* batches);
/* Timing of the batches' pricing kernels (enable = 1; off by default): each batch launched while
 * it is on brackets its kernel with HIP timing events on its own stream, giving an interval
 * [begin, end] on the device clock for rb_server_report_get (kernel time, union, overlap). */
int rb_server_set_timing(rb_server* s, int enable);
/* What the server has done so far.  The kernel fields cover the batches launched with timing on
 * (zero otherwise): their sum, union, time covered by two or more and the most at one instant. */
typedef struct rb_server_report {
    double p50_ms, p99_ms;     /* request latency percentiles (submit to results copied)        */
    long completed;            /* requests completed                                           */
    double path_steps;         /* sum of M*N over them                                         */
    long batches;              /* batches launched                                             */
    long peak_inflight;        /* most batches launched and not yet retired at once (host)     */
    long kernel_launches;      /* pricing-kernel intervals recorded                            */
    double kernel_ms;          /* sum of the pricing-kernel intervals                          */
    double device_busy_ms;     /* their union                                                  */
    double device_overlap_ms;  /* time covered by two or more of them                          */
    double device_span_ms;     /* first begin to last end                                      */
    long max_concurrent;       /* most intervals covering one instant                          */
    double conv_flops;         /* algorithmic Toeplitz flops M N (N-1) per distinct H, summed  */
} rb_server_report;
int rb_server_report_get(rb_server* s, rb_server_report* out);
void rb_server_destroy(rb_server* s);

/* Implied vols to the reference's tolerance, on the device: IV_call of rBergLib/BlackScholes.cpp:
 * 36-92 (S0 = 1, r = 0; K < 1 solved on the put by put-call parity) with Boost's bracketing from
 * the guess 0.2 by the factor 2 (at most 20 evaluations) and the bracket's midpoint once its
 * relative width is 2^-10 (eps_tolerance(11)).  Boost narrows the bracket with TOMS 748, this
 * routine by halving it, so the values are those the reference returns to within its tolerance,
 * not digit for digit.  NaN where the reference's solver throws.  The pricing entry points solve
 * to 1e-14 relative instead; this one reproduces what the reference's callers see. */
int rb_iv_reference(const double* price, const double* K, const double* T, int n, double* iv);

/* ---- pricing without device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): rb_price_v2 on the
 * fused kernel (cfg->precision RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM,
 * hybrid scheme, N <= 1024) with every device buffer beyond the caller's host inputs and
 * outputs -- tables, work items, per-item option sums, the queue counter, device results --
 * carved from a static per-device scratch array of 64 MiB that the library holds from load until
 * the process ends; it calls no hipMalloc, hipFree or other allocating routine.  A request whose
 * buffers exceed the scratch (M * out_size above about 4e9: the per-item FP64 option sums take
 * out_size * M / 64 bytes) is refused with RB_ERR_LIMIT.  Prices, implied vols and standard errors
 * are bit-identical to rb_price_v2 with the same cfg.  The call runs on `stream` (NULL: the null
 * stream) and returns when the results are on the host.  All calls on a device share its scratch,
 * which keeps the last call's contents; calls are serialised within the process (calls that
 * overlapped on one device would race on it). */
int rb_price_noalloc(double xi, const double* H, const double* eta, const double* rho, int par_size,
                     const double* T, const double* K, int op_size, int N, long M,
                     const uint64_t* seed, int seed_size, const rb_config* cfg, void* stream,
                     double* price, double* iv, double* stat, int out_size);

/* Host-only planning (no device call): group index of every option I and the unit of every
 * group, as the executor deduplicates them.  group_of_option has out_size entries,
 * unit_of_group at least out_size entries. */
int rb_layout(const double* H, const double* eta, const double* rho, int par_size,
              const double* T, const double* K, int op_size, int N, int method,
              int* nunits, int* ngroups, int* group_of_option, int* unit_of_group);

/* ---- probes used by the parity tests (device work, FP64) ---- */
/* Host only: Sobol' points i0 .. i0+n-1 of the quasi-Monte Carlo source in dims dimensions,
 * u[p*dims + d] (the generator behind rb_price_sobol). */
int rb_probe_sobol(int dims, long i0, long n, double* u);
/* The device Brownian-bridge mapping of rb_price_sobol_bridge on rows of N = 2^m values:
 * dW[r*N + i] from Z[r*N ..]. */
int rb_probe_bridge(const double* Z, long rows, int N, double* dW);
/* The device inverse normal CDF of the quasi-Monte Carlo source: z[i] = Phi^-1(u[i]). */
int rb_probe_norm_inv(const double* u, long n, double* z);
/* Host only: the exact scheme's covariance blocks A21[i][j] = Cov(Wtilde_i, W1_j),
 * A22[i][k] = Cov(Wtilde_i, Wtilde_k) (N x N each) and its factor L (N rows of 2N); any may be
 * NULL. */
int rb_probe_chol(double H, int N, double* A21, double* A22, double* L);
/* Philox4x32-10 on the device: ctr[4n], key[2] -> out[4n] */
int rb_probe_philox(const uint32_t* ctr, const uint32_t* key, long n, uint32_t* out);
/* Device Gaussians of paths [p0, p0+np): out[((p-p0)*nstreams + s)*N + j] */
int rb_probe_normals(const uint64_t* seed, int seed_size, long p0, long np, int nstreams, int N,
                     double* out);
/* Volterra process of every path of Z (2 blocks of N per path): Wtilde[m*N + i] */
int rb_probe_wtilde(const double* Z, long M, int N, double H, double* Wtilde);
/* A[32][16] by B[16][32] on the matrix cores with the operand layout of the fused kernel
 * (mode 0: one 32x32x16 MFMA on FP16 operands, 1: on BF16, 2: eight 32x32x2 FP32 MFMAs);
 * out[lane*16 + r] = accumulator register r of lane.  Mode 3: A[16][4] by B[4][16] (the first
 * 64 entries of each) on one 16x16x4 FP64 MFMA, out[lane*4 + r], as the FP64 convolution uses. */
int rb_probe_mfma_layout(const float* A, const float* B, int mode, float* out);
/* Device-to-device copy of `bytes` (multiple of 16) by a plain streaming kernel, reps times;
 * *avg_ms receives the mean kernel time (HIP events).  Used to calibrate HBM traffic counters. */
int rb_probe_copy(long bytes, int reps, double* avg_ms);
/* Issue cost of four instruction classes on the device: ns_per_inst[c] = nanoseconds per wave64
 * instruction per SIMD with every SIMD issuing one class back to back (eight independent chains
 * per lane, eight waves per SIMD); *simds (if not null) = the device's SIMDs (4 per CU).  c = 0:
 * FP32 FMA, 1: FP32 exp2 (transcendental), 2: 64-bit integer multiply-add of two 32-bit words
 * followed by a 32-bit add, 3: FP64 FMA.  No counterpart in the reference: the VALU roofline of
 * the fused kernels is priced with these. */
int rb_probe_issue(double* ns_per_inst, int* simds);
/* The split-precision kernels' per-path payoff for given (ln S', sigma, K) -> out[i], and its
 * Mills-ratio approximation at z[i] -> mills[i] (n values).  form 0: the fused GEMM kernel's
 * (deep out-of-the-money tail through erfcx), 1: the long-path kernel's (through erfc). */
int rb_probe_payoff(const double* lnS, const double* sig, const double* K, int n, double* out,
                    const float* z, float* mills, int form);
/* Per-path integrals (Ivdt, IsvdW) of option I from Z: out[2*(m*out_size + I) + {0,1}] */
int rb_probe_integrals(double xi, const double* H, const double* eta, const double* rho,
                       int par_size, const double* T, const double* K, int op_size, int N,
                       long M, const double* Z, const rb_config* cfg, double* out);

#ifdef __cplusplus
}
#endif

#endif /* RBERGOMI_MI355X_H */
