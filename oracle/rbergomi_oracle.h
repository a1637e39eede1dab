/*
 * rbergomi_oracle.h -- CPU restatement of the rough-Bergomi Romano-Touzi pricing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path
 * (rbergomi_amd/csrc) and the timed CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * library never links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain FP64 C++ and cites the
 * reference file:line it follows (paths relative to the reference repository root).
 *
 * Parity pins (what checks each function against something other than itself):
 *   orc_gamma            closed form, telescoping sum, H = 1/2 => Gamma_k = 1        (tests/test_oracle_math.py)
 *   orc_convolve_*       brute force in exact rationals, the reference's own known answer
 *                        x={1,2.4,3.9,4.8}, y={9,5.4,2.9,7.8} -> {9,27,50.96,79.02}
 *                        (rBergomi/text/tests.cpp:333-361), FFT == direct              (test_oracle_math.py)
 *   orc_wtilde           impulse responses (unit W1 / W1perp vectors) against the hybrid-scheme
 *                        weights written independently in the test; variance closed form (test_oracle_math.py)
 *   orc_payoff_rt_samples  the reference's own 200 single-thread per-path payoffs
 *                        rBergomi/testing/debug8.txt~ (tests/golden/debug8_payoffs.json);
 *                        eta = rho = 0 closed form; eta = 0 closed form; K -> 0 gives S' (test_oracle_golden.py)
 *   orc_compute_rt       reference known prices testRBergLib/tests.cpp:16-28 within 0.005 + stat;
 *                        E[S'] = 1 martingale identity within 4.5 SE                 (test_oracle_golden.py)
 *   orc_compute_mc       same known prices (the reference tests both methods against them)
 *   orc_iv_call          round trip IV(BS(sigma)) = sigma; put-call parity branch     (test_oracle_math.py)
 *   orc_iv_call_reference  within 2^-11 relative of the exact IV (the tolerance's guarantee);
 *                        a price of BS(0.2 * 2^k) gives 0.2 * 2^k exactly (a root on the bracket);
 *                        vols that the 20 evaluations cannot bracket give NaN, as Boost throws;
 *                        the parity branch                                              (test_oracle_math.py)
 *   orc_philox4x32_10    Random123 known-answer vectors                                (test_oracle_math.py)
 *   orc_normals_philox   distribution moments; equals Box-Muller on orc_philox4x32_10 outputs
 *   orc_turbo_rows       unbiased against the plain RT oracle (within 4.5 combined SE); eta = 0
 *                        closed form C_BS(1, K, T, sqrt(xi)); the control-variate mean equals its
 *                        closed-form adjustment; variance below the plain estimator's   (test_oracle_rows.py)
 *   orc_hyp2f1           scipy.special.hyp2f1 over the (a, b, c, z) the covariance uses    (test_chol_cpu.py)
 *   orc_chol_cov         Var Wtilde_t = t^(2H); H = 1/2 gives Brownian motion (min(s, t));
 *                        quadrature of 2H int (t-u)^(H-1/2) (s-u)^(H-1/2) du             (test_chol_cpu.py)
 *   orc_chol_factor      L L^T = A                                                      (test_chol_cpu.py)
 *   orc_payoff_rows_samples  equals orc_payoff_rt_samples on the zipped tensor laid out as rows;
 *                        per-row xi equals separate scalar-xi runs; row-order independent
 *                        (test_oracle_rows.py)
 *   orc_bridge           orthogonal map (unit vectors -> orthonormal rows); sum of dW = 2^(m/2) Z_0;
 *                        partial sums at dyadic points use only coarser coefficients and give the
 *                        Brownian-bridge midpoint law; m = 1, 2 written out                (test_qmc_cpu.py)
 *
 * Parity unpinned: the exact roundoff of FFTW (the reference convolution) and the iterate
 * sequence of Boost's bracket_and_solve_root in IV_call (rBergLib/BlackScholes.cpp:36-63).
 * Neither library exists in this environment; the direct-sum definition and the IV
 * round trip pin the mathematics instead.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* compensator of the variance process: 0 = reference form 0.5*eta^2*((i-1)*dt)^(2H)
 * (rBergLib/rBergomiMT.cpp:77-79), 1 = exact form 0.5*eta^2*(i*dt)^(2H)
 * (clean_Rfiles/hybridScheme.R:110-113). */
enum { ORC_COMP_REF = 0, ORC_COMP_EXACT = 1 };
/* convolution used inside orc_wtilde: 0 = direct sum, 1 = own radix-2 FFT */
enum { ORC_CONV_DIRECT = 0, ORC_CONV_FFT = 1 };

/* Gamma_0 = 0, Gamma_k = ((k+1)^(a+1) - k^(a+1))/(a+1), a = H - 1/2.
 * rBergLib/rBergomiMT.cpp:143-149 */
void orc_gamma(double H, int N, double* gamma);

/* truncated linear convolution z_i = sum_{k=0}^{i} x_k y_{i-k}, i < n.  This is what the
 * length-(2N-1) FFT product in updateWtilde returns in its first N entries
 * (rBergLib/rBergomiMT.cpp:91-105). */
void orc_convolve_direct(const double* x, const double* y, int n, double* z);
void orc_convolve_fft(const double* x, const double* y, int n, double* z);

/* Volterra process on the unit interval, value at t=(i+1)/N.
 * rBergLib/rBergomiMT.cpp:83-109 (single thread: rBergLib/RBergomi.cpp:397-423) */
void orc_wtilde(const double* W1, const double* W1perp, double H, int N, int conv_mode,
                double* Wtilde);

/* v_0 = xi, v_i = xi*exp(eta*T^H*Wtilde_{i-1} - compensator_i), dt = T/N.
 * scaleWtilde rBergLib/rBergomiMT.cpp:163-167, updateV :73-80 */
void orc_update_v(const double* Wtilde, double xi, double H, double eta, double T, int N,
                  int comp_mode, double* v);

/* intVdt / intRootVdW, rBergLib/rBergomiMT.cpp:427-436 */
double orc_int_vdt(const double* v, int N, double dt);
double orc_int_rootvdw(const double* v, const double* W1, int N, double sdt);

/* Phi and Black-Scholes, rBergLib/BlackScholes.cpp:16-33 */
double orc_phi(double x);
double orc_bs_call(double S0, double K, double tau, double sigma);
double orc_bs_put(double S0, double K, double tau, double sigma);

/* implied vol of a call (S0 > K handled through the put, as rBergLib/BlackScholes.cpp:36-42).
 * Safeguarded Newton to |dsigma| <= 1e-14*sigma instead of Boost's 2^-10 bracket.
 * Returns NaN when the price is outside the no-arbitrage range. */
double orc_iv_call(double price, double S0, double K, double tau);

/* implied vol to the reference's tolerance: Boost's bracketing from 0.2 by the factor 2 (at most
 * 20 evaluations), the midpoint of a bracket of relative width 2^-10 (rBergLib/BlackScholes.cpp:
 * 36-92); the bracket is halved where Boost runs TOMS 748.  NaN where Boost throws. */
double orc_iv_call_reference(double price, double S0, double K, double tau);

/* Romano-Touzi conditional payoff of one path for one option, given the path integrals.
 * rBergLib/rBergomiMT.cpp:383-388 */
double orc_rt_payoff(double Ivdt, double IsvdW, double rho, double K);

/* Per-path Romano-Touzi payoffs from user Gaussians.
 * Z layout (rBergLib/cinterface.cpp:113-119): path m uses W1 = Z[2mN, 2mN+N),
 * W1perp = Z[2mN+N, 2mN+2N).  payoff[m*out_size + I], I = h*op_size + t
 * (rBergLib/ParamTot.h:32-33).  Loop and trigger structure of
 * ComputePayoffRTsamples_ST, rBergLib/rBergomiMT.cpp:528-610. */
void orc_payoff_rt_samples(double xi, const double* H, const double* eta, const double* rho,
                           int par_size, const double* T, const double* K, int op_size,
                           int N, long M, const double* Z, int comp_mode, double* payoff);

/* Per-path integrals (Ivdt, IsvdW) for option I from the same Z layout: out[2*(m*out_size+I)+{0,1}] */
void orc_integrals_rt_samples(double xi, const double* H, const double* eta, const double* rho,
                              int par_size, const double* T, const double* K, int op_size,
                              int N, long M, const double* Z, int comp_mode, double* out);

/* Per-path plain Monte Carlo payoffs (S_T - K)^+ with three Gaussian blocks per path:
 * W1 = Z[3mN, 3mN+N), W1perp = Z[3mN+N, 3mN+2N), Wperp = Z[3mN+2N, 3mN+3N).
 * ComputePriceMT loop body, rBergLib/rBergomiMT.cpp:231-267 (updateZ :134-136,
 * scaleZ :129-132, updateS :111-116). */
void orc_payoff_mc_samples(double xi, const double* H, const double* eta, const double* rho,
                           int par_size, const double* T, const double* K, int op_size,
                           int N, long M, const double* Z, int comp_mode, double* payoff);

/* Per-path Romano-Touzi payoffs of unordered rows: row i prices (T[i], K[i]) under
 * (xi[i], H[i], eta[i], rho[i]), every row on the same W1, W1perp of the path (same Z layout as
 * orc_payoff_rt_samples); payoff[m*rows + i].  updatePayoffUnordered, rBergomi/rBergomiMT.cpp:
 * 939-960, inside the path loop of ComputePriceRTMTunstructured :837-911 (its xi is one scalar;
 * the per-row xi follows ML_console_ST/ParamTot.h:25-43).  Direct-sum convolution. */
void orc_payoff_rows_samples(const double* xi, const double* H, const double* eta,
                             const double* rho, const double* T, const double* K, int rows,
                             int N, long M, const double* Z, int comp_mode, double* payoff);

/* Turbo-charged Romano-Touzi estimator of unordered rows (ComputePriceRTVarRed,
 * ML_console_ST/RBergomi.cpp:219-288; estimateControlVariate :493-537;
 * updatePayoffControlVariate :478-491; covariance ML_console_ST/aux.h:207-220):
 *   pilot (Mc paths of Zc): Q_i = max 2 Iv; X = C_BS(S', K, 1, sqrt((1-rho^2) Iv)),
 *   Y = C_BS(S', K, 1, sqrt(rho^2 (Q - Iv))); alpha_i = -cov(X, Y) / cov(Y, Y) (two-pass sample
 *   covariances; alpha = 0 when cov(Y, Y) = 0, where the reference divides by zero);
 *   main (M paths of Z): payoff = (X + alpha Y)(W)/2 + (X + alpha Y)(-W)/2 (antithetic), with
 *   qmax in place of Q when Q <= Iv, counted per evaluation; price = mean - alpha [(1 - p)
 *   C_BS(1, K, 1, sqrt(rho^2 Q)) + p C_BS(1, K, 1, sqrt(rho^2 qmax))], p = count / M as in the
 *   reference; stat from the pair-averaged payoffs.  Z layouts as orc_payoff_rt_samples.
 *   exceptions[i] = count. */
void orc_turbo_rows(const double* xi, const double* H, const double* eta, const double* rho,
                    const double* T, const double* K, int rows, int N, long Mc, const double* Zc,
                    long M, const double* Z, double qmax, int comp_mode, double* price,
                    double* stat, double* alpha, double* Q, long* exceptions);

/* Gauss hypergeometric 2F1(a, b; c; z) for 0 <= z <= 1 with c - a - b > 0 non-integer: the Gauss
 * series, and for z > 1/2 the transformation to 1 - z (Abramowitz & Stegun 15.3.6).  Stands in
 * for gsl_sf_hyperg_2F1 (GSL is absent), used by RfBm::G, PremiaRBergomi/RfBm.cpp:31-37. */
double orc_hyp2f1(double a, double b, double c, double z);

/* The exact joint covariance of (W1_0..W1_{n-1}, Wtilde_0..Wtilde_{n-1}) as RfBm::A
 * (PremiaRBergomi/RfBm.cpp:18-58), A[2n][2n] row-major, and its full 2n x 2n Cholesky factor as
 * the RfBm constructor computes it (:63-80), L[2n][2n].  Returns 0, or -1 if A is not positive
 * definite. */
void orc_chol_cov(double H, int n, double* A);
int orc_chol_factor(double H, int n, double* L);

/* Per-path payoffs with Wtilde sampled exactly (RfBm::generate, :82-95): path m's first two
 * blocks x1, x2 of Z give W1 = x1 and Wtilde = L (x1, x2) (rows n..2n-1 of L); the rest as
 * orc_payoff_rt_samples (method 0) or orc_payoff_mc_samples (method 1, third block Wperp).
 * updatePayoff_cholesky, PremiaRBergomi/rBergomi.cpp:190-204, for the RT payoff. */
int orc_payoff_chol_samples(double xi, const double* H, const double* eta, const double* rho,
                            int par_size, const double* T, const double* K, int op_size, int N,
                            long M, const double* Z, int comp_mode, int method, double* payoff);

/* price/stat from a payoff matrix [M][out_size]: mean and sqrt(E[X^2]-mean^2)/sqrt(M),
 * rBergLib/rBergomiMT.cpp:404-409 */
void orc_price_stat(const double* payoff, long M, int out_size, double* price, double* stat);

/* Reference-structured seeded pricers (the timed CPU baseline).  Loop structure of
 * ComputePriceRTMT / ComputeIVRTMT (rBergLib/rBergomiMT.cpp:301-425): mt19937_64 +
 * std::normal_distribution per thread seeded through std::seed_seq (RNG::RNG :63-71;
 * single thread: RBergomiST::setGen rBergLib/RBergomi.cpp:381-389), per-path FFT
 * convolution, the options loop with the memoisation triggers, thread-private sums and a
 * critical-section reduction.  The W1perp data race of the reference (:348-349) is fixed:
 * W1perp is private per thread.  conv_mode selects FFT (reference structure) or direct.
 * iv may be NULL. */
void orc_compute_rt(double xi, const double* H, const double* eta, const double* rho,
                    int par_size, const double* T, const double* K, int op_size, int N, long M,
                    const uint64_t* seed, int seed_size, int num_threads, int conv_mode,
                    double* price, double* iv, double* stat);
/* plain Monte Carlo sibling, ComputePriceMT / ComputeIVMT rBergLib/rBergomiMT.cpp:174-299 */
void orc_compute_mc(double xi, const double* H, const double* eta, const double* rho,
                    int par_size, const double* T, const double* K, int op_size, int N, long M,
                    const uint64_t* seed, int seed_size, int num_threads, int conv_mode,
                    double* price, double* iv, double* stat);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based generator of the
 * product path.  ctr[4], key[2] -> out[4]. */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
/* 64-bit Philox key from the seed array (splitmix64 fold), as the product path derives it */
void orc_philox_key(const uint64_t* seed, int seed_size, uint32_t* key);
/* Standard normals of the product path: normal number j of stream s of path p is
 * Box-Muller on Philox(ctr = {j/4, p_lo, p_hi, s}); z[p][s][j] written to
 * out[(p*nstreams + s)*N + j] for p in [p0, p0+np). */
void orc_normals_philox(const uint64_t* seed, int seed_size, long p0, long np, int nstreams,
                        int N, double* out);
/* hierarchical2increments (rBergomi/aux.h:105-125): Z[2^m] (bridge coefficients) -> dW[2^m] */
void orc_bridge(const double* Z, int m, double* dW);

#ifdef __cplusplus
}
#endif
