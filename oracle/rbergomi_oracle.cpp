/*
 * rbergomi_oracle.cpp -- CPU restatement of the reference rough-Bergomi pricing path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The product library
 * under rbergomi_amd/ never links or calls this file.  See rbergomi_oracle.h for the list of
 * checks that pin each function.
 *
 * Reference semantics restated here (paths relative to the reference repository root):
 *   rBergLib/RBergomi.cpp:157-232, 381-491   single-thread class (the semantic oracle)
 *   rBergLib/rBergomiMT.cpp:63-172, 301-610  OpenMP drivers and numerics
 *   rBergLib/BlackScholes.cpp:16-92          Phi, BS prices, implied vol
 *   rBergLib/ParamTot.h:20-50                index algebra and memoisation triggers
 *   rBergLib/cinterface.cpp:49-131           C ABI layouts
 */
#include "rbergomi_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::vector<double> Vec;

/* ParamTot: (H, eta, rho) zipped into par_size triples, (T, K) zipped into op_size pairs,
 * I = h*op_size + t; a trigger fires when the value differs from the one at I-1 by more
 * than 1e-10.  rBergLib/ParamTot.h:20-50, rBergLib/ParamTot.cpp:18-29 */
struct Params {
    const double *H, *eta, *rho, *T, *K;
    long par_size, op_size;
    double eps = 1e-10;
    long size() const { return par_size * op_size; }
    double h(long I) const { return H[I / op_size]; }
    double e(long I) const { return eta[I / op_size]; }
    double r(long I) const { return rho[I / op_size]; }
    double t(long I) const { return T[I % op_size]; }
    double k(long I) const { return K[I % op_size]; }
    bool HTrig(long I) const { return I == 0 || std::fabs(h(I) - h(I - 1)) > eps; }
    bool etaTrig(long I) const { return I == 0 || std::fabs(e(I) - e(I - 1)) > eps; }
    bool rhoTrig(long I) const { return I == 0 || std::fabs(r(I) - r(I - 1)) > eps; }
    bool TTrig(long I) const { return I == 0 || std::fabs(t(I) - t(I - 1)) > eps; }
};

/* In-place iterative radix-2 FFT (Cooley-Tukey, bit reversal then butterflies); sign -1
 * forward, +1 backward (unnormalised).  n must be a power of two. */
void fft_pow2(std::complex<double>* a, int n, int sign) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / len;
        const std::complex<double> wl(std::cos(ang), std::sin(ang));
        for (int i = 0; i < n; i += len) {
            std::complex<double> w(1.0, 0.0);
            for (int j = 0; j < len / 2; ++j) {
                std::complex<double> u = a[i + j], v = a[i + j + len / 2] * w;
                a[i + j] = u + v;
                a[i + j + len / 2] = u - v;
                w *= wl;
            }
        }
    }
}

/* FFT workspace of one thread: the reference keeps three complex arrays and plans per
 * thread (fftData, rBergLib/rBergomiMT.cpp:34-61); here the length is the power of two
 * >= 2N-1 instead of 2N-1 (a linear convolution is exact for any length >= 2N-1). */
struct FFTWork {
    int n = 0;
    std::vector<std::complex<double>> x, y;
    explicit FFTWork(int N) {
        n = 1;
        while (n < 2 * N - 1) n <<= 1;
        x.resize(n);
        y.resize(n);
    }
    /* three transforms per call, as updateWtilde does (:93-102) */
    void convolve(const double* a, const double* b, int N, double* z) {
        for (int i = 0; i < n; ++i) {
            x[i] = i < N ? std::complex<double>(a[i], 0.0) : 0.0;
            y[i] = i < N ? std::complex<double>(b[i], 0.0) : 0.0;
        }
        fft_pow2(x.data(), n, -1);
        fft_pow2(y.data(), n, -1);
        for (int i = 0; i < n; ++i) x[i] *= y[i];
        fft_pow2(x.data(), n, +1);
        for (int i = 0; i < N; ++i) z[i] = x[i].real() / n;
    }
};

/* updateWtilde, rBergLib/rBergomiMT.cpp:83-109: W1hat from the 2x2 first-cell Cholesky,
 * Y2 = Gamma * W1 (truncated), Wtilde = sqrt(2H)*N^-H*(Y2 + W1hat). */
void wtilde_impl(const double* W1, const double* W1perp, const double* gamma, double H, int N,
                 FFTWork* fft, double* Wt) {
    const double s2H = std::sqrt(2.0 * H);
    const double rhoH = s2H / (H + 0.5);
    const double a = rhoH / s2H, b = std::sqrt(1.0 - rhoH * rhoH) / s2H;
    Vec Y2(N);
    if (fft)
        fft->convolve(W1, gamma, N, Y2.data());
    else
        orc_convolve_direct(W1, gamma, N, Y2.data());
    const double sc = s2H * std::pow(1.0 / N, H);
    for (int i = 0; i < N; ++i) Wt[i] = sc * Y2[i] + sc * (a * W1[i] + b * W1perp[i]);
}

/* scaleWtilde + updateV, rBergLib/rBergomiMT.cpp:163-167 and :73-80 */
void v_impl(const double* Wt, double xi, double H, double eta, double T, int N, int comp,
            double* v) {
    const double dt = T / N;
    const double TH = std::pow(T, H);
    v[0] = xi;
    for (int i = 1; i < N; ++i) {
        const double tc = comp == ORC_COMP_EXACT ? i * dt : (i - 1) * dt;
        v[i] = xi * std::exp(eta * (TH * Wt[i - 1]) - 0.5 * eta * eta * std::pow(tc, 2 * H));
    }
}

/* updateZ + scaleZ + updateS, rBergLib/rBergomiMT.cpp:111-136 */
double s_impl(const double* v, const double* W1, const double* Wperp, double rho, double dt,
              int N) {
    const double sdt = std::sqrt(dt), c = std::sqrt(1.0 - rho * rho);
    double X = 0.0;
    for (int i = 0; i < N; ++i) {
        const double Z = rho * W1[i] + c * Wperp[i];
        X += std::sqrt(v[i]) * (sdt * Z) - 0.5 * v[i] * dt;
    }
    return std::exp(X);
}

/* One path through the options loop with the memoisation triggers of
 * ComputePayoffRTsamples_ST (rBergLib/rBergomiMT.cpp:576-606).  Writes the payoff (and the
 * two integrals, if ints != nullptr) of every option I. */
struct PathState {
    Vec Wt, v;
    double dt = 0.0, sdt = 0.0;
    explicit PathState(int N) : Wt(N), v(N) {}
};

void rt_path(const Params& par, double xi, int N, const double* W1, const double* W1perp,
             int comp, FFTWork* fft, const std::vector<Vec>& gammas, PathState& st,
             double* payoff, double* ints) {
    for (long I = 0; I < par.size(); ++I) {
        bool update = par.HTrig(I);
        if (update) wtilde_impl(W1, W1perp, gammas[I / par.op_size].data(), par.h(I), N, fft, st.Wt.data());
        update = update || par.TTrig(I);
        if (update) {
            st.dt = par.t(I) / N;
            st.sdt = std::sqrt(st.dt);
        }
        update = update || par.etaTrig(I);
        if (update) v_impl(st.Wt.data(), xi, par.h(I), par.e(I), par.t(I), N, comp, st.v.data());
        const double Iv = orc_int_vdt(st.v.data(), N, st.dt);
        const double Is = orc_int_rootvdw(st.v.data(), W1, N, st.sdt);
        payoff[I] = orc_rt_payoff(Iv, Is, par.r(I), par.k(I));
        if (ints) {
            ints[2 * I] = Iv;
            ints[2 * I + 1] = Is;
        }
    }
}

std::vector<Vec> gamma_tables(const Params& par, int N) {
    std::vector<Vec> g(par.par_size, Vec(N));
    for (long h = 0; h < par.par_size; ++h) orc_gamma(par.H[h], N, g[h].data());
    return g;
}

double posPart(double x) { return x > 0 ? x : 0; }

/* Black-Scholes price as a function of sigma for the root solver */
double bs_otm(double S0, double K, double tau, double sigma, bool call) {
    return call ? orc_bs_call(S0, K, tau, sigma) : orc_bs_put(S0, K, tau, sigma);
}

/* 64-bit finaliser of splitmix64 (Steele, Lea, Flood, OOPSLA'14) */
uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

void orc_gamma(double H, int N, double* gamma) {
    const double alpha = H - 0.5;
    if (N <= 0) return;
    gamma[0] = 0.0;
    for (int i = 1; i < N; ++i)
        gamma[i] = (std::pow(i + 1.0, alpha + 1.0) - std::pow(double(i), alpha + 1.0)) / (alpha + 1.0);
}

void orc_convolve_direct(const double* x, const double* y, int n, double* z) {
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k <= i; ++k) s += x[k] * y[i - k];
        z[i] = s;
    }
}

void orc_convolve_fft(const double* x, const double* y, int n, double* z) {
    FFTWork w(n);
    w.convolve(x, y, n, z);
}

void orc_wtilde(const double* W1, const double* W1perp, double H, int N, int conv_mode,
                double* Wtilde) {
    Vec g(N);
    orc_gamma(H, N, g.data());
    if (conv_mode == ORC_CONV_FFT) {
        FFTWork w(N);
        wtilde_impl(W1, W1perp, g.data(), H, N, &w, Wtilde);
    } else {
        wtilde_impl(W1, W1perp, g.data(), H, N, nullptr, Wtilde);
    }
}

void orc_update_v(const double* Wtilde, double xi, double H, double eta, double T, int N,
                  int comp_mode, double* v) {
    v_impl(Wtilde, xi, H, eta, T, N, comp_mode, v);
}

double orc_int_vdt(const double* v, int N, double dt) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += v[i];
    return dt * s;
}

double orc_int_rootvdw(const double* v, const double* W1, int N, double sdt) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += std::sqrt(v[i]) * sdt * W1[i];
    return s;
}

double orc_phi(double x) { return 0.5 * std::erfc(-x * M_SQRT1_2); }

double orc_bs_call(double S0, double K, double tau, double sigma) {
    const double d1 = (std::log(S0 / K) + 0.5 * sigma * sigma * tau) / (sigma * std::sqrt(tau));
    const double d2 = d1 - sigma * std::sqrt(tau);
    return orc_phi(d1) * S0 - orc_phi(d2) * K;
}

double orc_bs_put(double S0, double K, double tau, double sigma) {
    const double d1 = (std::log(S0 / K) + 0.5 * sigma * sigma * tau) / (sigma * std::sqrt(tau));
    const double d2 = d1 - sigma * std::sqrt(tau);
    return -orc_phi(-d1) * S0 + orc_phi(-d2) * K;
}

double orc_iv_call(double price, double S0, double K, double tau) {
    /* S0 > K: solve on the out-of-the-money put through put-call parity
     * (rBergLib/BlackScholes.cpp:36-40, r = 0) */
    bool call = true;
    double target = price;
    if (S0 > K) {
        call = false;
        target = price - S0 + K;
    }
    const double upper = call ? S0 : K; /* sup over sigma of the OTM price */
    if (!(target > 0.0) || !(target < upper) || !(tau > 0.0)) return std::nan("");
    double lo = 0.0, hi = 1.0;
    while (bs_otm(S0, K, tau, hi, call) < target) {
        lo = hi;
        hi *= 2.0;
        if (hi > 1e6) return std::nan("");
    }
    double s = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
        const double f = bs_otm(S0, K, tau, s, call) - target;
        if (f > 0) hi = s; else lo = s;
        const double d1 = (std::log(S0 / K) + 0.5 * s * s * tau) / (s * std::sqrt(tau));
        const double vega = S0 * std::exp(-0.5 * d1 * d1) / std::sqrt(2.0 * M_PI) * std::sqrt(tau);
        double sn = vega > 0 ? s - f / vega : 0.5 * (lo + hi);
        if (!(sn > lo && sn < hi)) sn = 0.5 * (lo + hi);
        const double ds = std::fabs(sn - s);
        s = sn;
        if (ds <= 1e-14 * s || hi - lo <= 1e-15 * s) break;
    }
    return s;
}

double orc_iv_call_reference(double price, double S0, double K, double tau) {
    /* IV_call / IV_put of rBergLib/BlackScholes.cpp:36-92 to the reference's tolerance.  S0 > K:
     * the put through put-call parity (:38-41, r = 0).  Boost's bracket_and_solve_root on
     * f(sigma) = target - BS(sigma) (decreasing, is_rising = false, :47-60): from guess = 0.2 the
     * bracket walks by factor = 2 -- down while f < 0, up while f >= 0 -- for at most maxit = 20
     * evaluations, then narrows until |b - a| <= 2^(1 - 11) min(|a|, |b|) (eps_tolerance(11)),
     * and the midpoint a + (b - a)/2 is returned (:63).  Boost narrows with TOMS 748 (Alefeld,
     * Potra and Shi 1995); this restatement halves the bracket: the result lies within the same
     * tolerance of the root, not on Boost's iterates.  NaN where Boost throws. */
    bool call = true;
    double target = price;
    if (S0 > K) {
        call = false;
        target = price - S0 + K;
    }
    const double upper = call ? S0 : K;
    if (!(target > 0.0) || !(target < upper) || !(tau > 0.0)) return std::nan("");
    auto f = [&](double s) { return target - (call ? orc_bs_call(S0, K, tau, s) : orc_bs_put(S0, K, tau, s)); };
    double a = 0.2, b = 0.2, fa = f(a), fb = fa;
    int count = 20 - 1;
    if (fa < 0.0) {
        while ((fa < 0.0) == (fb < 0.0)) {
            if (count-- == 0) return std::nan("");
            b = a;
            fb = fa;
            a /= 2.0;
            fa = f(a);
        }
    } else {
        while ((fa < 0.0) == (fb < 0.0)) {
            if (count-- == 0) return std::nan("");
            a = b;
            fa = fb;
            b *= 2.0;
            fb = f(b);
        }
    }
    for (int it = 0; it < 200 && !(std::fabs(b - a) <= std::ldexp(1.0, -10) * std::fmin(std::fabs(a), std::fabs(b)));
         ++it) {
        if (fa == 0.0) return a;
        if (fb == 0.0) return b;
        const double m = a + 0.5 * (b - a), fm = f(m);
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

double orc_rt_payoff(double Ivdt, double IsvdW, double rho, double K) {
    const double vol = std::sqrt((1.0 - rho * rho) * Ivdt);
    const double spot = std::exp(-0.5 * rho * rho * Ivdt + rho * IsvdW);
    return orc_bs_call(spot, K, 1.0, vol);
}

void orc_payoff_rt_samples(double xi, const double* H, const double* eta, const double* rho,
                           int par_size, const double* T, const double* K, int op_size,
                           int N, long M, const double* Z, int comp_mode, double* payoff) {
    Params par{H, eta, rho, T, K, par_size, op_size};
    const std::vector<Vec> g = gamma_tables(par, N);
    const long out = par.size();
#pragma omp parallel
    {
        PathState st(N);
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m)
            rt_path(par, xi, N, Z + 2 * m * N, Z + (2 * m + 1) * N, comp_mode, nullptr, g, st,
                    payoff + m * out, nullptr);
    }
}

void orc_integrals_rt_samples(double xi, const double* H, const double* eta, const double* rho,
                              int par_size, const double* T, const double* K, int op_size,
                              int N, long M, const double* Z, int comp_mode, double* out) {
    Params par{H, eta, rho, T, K, par_size, op_size};
    const std::vector<Vec> g = gamma_tables(par, N);
    const long os = par.size();
#pragma omp parallel
    {
        PathState st(N);
        Vec pay(os);
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m)
            rt_path(par, xi, N, Z + 2 * m * N, Z + (2 * m + 1) * N, comp_mode, nullptr, g, st,
                    pay.data(), out + 2 * m * os);
    }
}

void orc_payoff_rows_samples(const double* xi, const double* H, const double* eta,
                             const double* rho, const double* T, const double* K, int rows,
                             int N, long M, const double* Z, int comp_mode, double* payoff) {
#pragma omp parallel
    {
        Vec Wt(N), v(N);
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m) {
            const double* W1 = Z + 2 * m * N;
            const double* W1p = W1 + N;
            for (int i = 0; i < rows; ++i) {
                // updateWtilde, scaleWtilde + updateV, intVdt, intRootVdW, BS_call_price
                orc_wtilde(W1, W1p, H[i], N, ORC_CONV_DIRECT, Wt.data());
                orc_update_v(Wt.data(), xi[i], H[i], eta[i], T[i], N, comp_mode, v.data());
                const double dt = T[i] / N;
                const double Iv = orc_int_vdt(v.data(), N, dt);
                const double Is = orc_int_rootvdw(v.data(), W1, N, std::sqrt(dt));
                payoff[m * rows + i] = orc_rt_payoff(Iv, Is, rho[i], K[i]);
            }
        }
    }
}

// integrals of row i on one path (updateWtilde, scaleWtilde + updateV, intVdt, intRootVdW)
static void row_integrals(const double* W1, const double* W1p, double xi, double H, double eta,
                          double T, int N, int comp_mode, Vec& Wt, Vec& v, double& Iv, double& Is) {
    orc_wtilde(W1, W1p, H, N, ORC_CONV_DIRECT, Wt.data());
    orc_update_v(Wt.data(), xi, H, eta, T, N, comp_mode, v.data());
    const double dt = T / N;
    Iv = orc_int_vdt(v.data(), N, dt);
    Is = orc_int_rootvdw(v.data(), W1, N, std::sqrt(dt));
}

// two-pass sample covariance, ML_console_ST/aux.h:207-220
static double sample_cov(const Vec& x, const Vec& y) {
    double mx = 0.0, my = 0.0;
    for (double a : x) mx += a;
    for (double b : y) my += b;
    mx /= double(x.size());
    my /= double(y.size());
    double c = 0.0;
    for (size_t k = 0; k < x.size(); ++k) c += (x[k] - mx) * (y[k] - my);
    return c / double(x.size() - 1);
}

void orc_turbo_rows(const double* xi, const double* H, const double* eta, const double* rho,
                    const double* T, const double* K, int rows, int N, long Mc, const double* Zc,
                    long M, const double* Z, double qmax, int comp_mode, double* price,
                    double* stat, double* alpha, double* Q, long* exceptions) {
    Vec Wt(N), v(N), mW1(N), mW1p(N);
    // pilot: integrals, Q = max 2 Iv (estimateControlVariate :504-516)
    std::vector<Vec> Iv(rows, Vec(Mc)), Is(rows, Vec(Mc));
    for (int i = 0; i < rows; ++i) Q[i] = 0.0;
    for (long m = 0; m < Mc; ++m) {
        const double* W1 = Zc + 2 * m * N;
        for (int i = 0; i < rows; ++i) {
            row_integrals(W1, W1 + N, xi[i], H[i], eta[i], T[i], N, comp_mode, Wt, v, Iv[i][m], Is[i][m]);
            Q[i] = std::max(Q[i], 2.0 * Iv[i][m]);
        }
    }
    // X, Y and alpha = -cov(X, Y) / cov(Y, Y)  (:520-534)
    for (int i = 0; i < rows; ++i) {
        Vec X(Mc), Y(Mc);
        const double r2 = rho[i] * rho[i];
        for (long m = 0; m < Mc; ++m) {
            const double S = std::exp(-0.5 * r2 * Iv[i][m] + rho[i] * Is[i][m]);
            X[m] = orc_bs_call(S, K[i], 1.0, std::sqrt((1.0 - r2) * Iv[i][m]));
            Y[m] = orc_bs_call(S, K[i], 1.0, std::sqrt(r2 * (Q[i] - Iv[i][m])));
        }
        const double vy = sample_cov(Y, Y);
        alpha[i] = vy > 0.0 ? -sample_cov(X, Y) / vy : 0.0;
    }
    // main loop with antithetics (:245-266) and the control-variate payoff (:478-491)
    std::vector<double> s1(rows, 0.0), s2(rows, 0.0);
    for (int i = 0; i < rows; ++i) exceptions[i] = 0;
    for (long m = 0; m < M; ++m) {
        const double* W1 = Z + 2 * m * N;
        for (int k = 0; k < N; ++k) {
            mW1[k] = -W1[k];
            mW1p[k] = -W1[N + k];
        }
        for (int i = 0; i < rows; ++i) {
            const double r2 = rho[i] * rho[i];
            double pay = 0.0;
            for (int side = 0; side < 2; ++side) {
                double iv_, is_;
                if (side == 0)
                    row_integrals(W1, W1 + N, xi[i], H[i], eta[i], T[i], N, comp_mode, Wt, v, iv_, is_);
                else
                    row_integrals(mW1.data(), mW1p.data(), xi[i], H[i], eta[i], T[i], N, comp_mode, Wt, v,
                                  iv_, is_);
                const double S = std::exp(-0.5 * r2 * iv_ + rho[i] * is_);
                const double X = orc_bs_call(S, K[i], 1.0, std::sqrt((1.0 - r2) * iv_));
                const bool exc = !(Q[i] > iv_);
                if (exc) ++exceptions[i];
                const double Y = orc_bs_call(S, K[i], 1.0, std::sqrt(r2 * ((exc ? qmax : Q[i]) - iv_)));
                pay += 0.5 * (X + alpha[i] * Y);
            }
            s1[i] += pay;
            s2[i] += pay * pay;
        }
    }
    for (int i = 0; i < rows; ++i) {
        const double mean = s1[i] / double(M);
        const double var = s2[i] / double(M) - mean * mean;
        stat[i] = std::sqrt(var) / std::sqrt(double(M));
        const double pq = double(exceptions[i]) / double(M);
        const double r2 = rho[i] * rho[i];
        const double adj = (1.0 - pq) * orc_bs_call(1.0, K[i], 1.0, std::sqrt(r2 * Q[i])) +
                           pq * orc_bs_call(1.0, K[i], 1.0, std::sqrt(r2 * qmax));
        price[i] = mean - alpha[i] * adj;
    }
}

static double orc_gauss_series(double a, double b, double c, double z) {
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 2000; ++k) {
        term *= (a + k) * (b + k) / ((c + k) * (k + 1.0)) * z;
        sum += term;
        if (std::fabs(term) < 1e-18 * std::fabs(sum)) break;
    }
    return sum;
}

double orc_hyp2f1(double a, double b, double c, double z) {
    if (z <= 0.5) return orc_gauss_series(a, b, c, z);
    const double w = 1.0 - z, s = c - a - b;
    return std::tgamma(c) * std::tgamma(s) / (std::tgamma(c - a) * std::tgamma(c - b)) *
               orc_gauss_series(a, b, 1.0 - s, w) +
           std::pow(w, s) * std::tgamma(c) * std::tgamma(-s) / (std::tgamma(a) * std::tgamma(b)) *
               orc_gauss_series(c - a, c - b, 1.0 + s, w);
}

// RfBm::A / G / cBm / cfBm / cMixed, PremiaRBergomi/RfBm.cpp:18-58
static double rfbm_G(double H, double x) {
    const double g = 0.5 - H;
    if (g == 0.0) return 1.0;
    return 2.0 * H * (std::pow(x, -g) / (1.0 - g) +
                      g * std::pow(x, -(1.0 + g)) * orc_hyp2f1(1.0, 1.0 + g, 3.0 - g, 1.0 / x) /
                          ((1.0 - g) * (2.0 - g)));
}

static double rfbm_A(double H, int n, int i, int j) {
    if (i < n && j < n) return i == j ? 1.0 : 0.0;
    if (i >= n && j >= n) {
        const double t = double(std::min(i, j) - n + 1) / n, s = double(std::max(i, j) - n + 1) / n;
        return std::pow(t, 2 * H) * rfbm_G(H, s / t);
    }
    const int b = i < n ? i : j, f = i < n ? j - n : i - n;  // Bm index b, fBm index f
    if (b > f) return 0.0;
    const double u = double(b) / n, v = double(b + 1) / n, t = double(f + 1) / n;
    return 2.0 * std::sqrt(2.0 * H * n) * (std::pow(t - u, 0.5 + H) - std::pow(t - v, 0.5 + H)) /
           (1.0 + 2.0 * H);
}

void orc_chol_cov(double H, int n, double* A) {
    for (int i = 0; i < 2 * n; ++i)
        for (int j = 0; j < 2 * n; ++j) A[size_t(i) * 2 * n + j] = rfbm_A(H, n, i, j);
}

int orc_chol_factor(double H, int n, double* L) {
    const int m = 2 * n;
    std::vector<double> A(size_t(m) * m);
    orc_chol_cov(H, n, A.data());
    for (size_t k = 0; k < size_t(m) * m; ++k) L[k] = 0.0;
    for (int k = 0; k < m; ++k) {  // column-by-column as the RfBm constructor
        double temp = 0.0;
        for (int j = 0; j < k; ++j) temp += L[size_t(k) * m + j] * L[size_t(k) * m + j];
        const double d = A[size_t(k) * m + k] - temp;
        if (!(d > 0.0)) return -1;
        L[size_t(k) * m + k] = std::sqrt(d);
        for (int i = k + 1; i < m; ++i) {
            double t2 = 0.0;
            for (int j = 0; j < k; ++j) t2 += L[size_t(i) * m + j] * L[size_t(k) * m + j];
            L[size_t(i) * m + k] = (A[size_t(i) * m + k] - t2) / L[size_t(k) * m + k];
        }
    }
    return 0;
}

int orc_payoff_chol_samples(double xi, const double* H, const double* eta, const double* rho,
                            int par_size, const double* T, const double* K, int op_size, int N,
                            long M, const double* Z, int comp_mode, int method, double* payoff) {
    const int ns = method == 1 ? 3 : 2;
    const long out = long(par_size) * op_size;
    // one factor per parameter set (the reference builds one RfBm per H)
    std::vector<std::vector<double>> Ls(par_size, std::vector<double>(size_t(4) * N * N));
    for (int h = 0; h < par_size; ++h)
        if (orc_chol_factor(H[h], N, Ls[h].data())) return -1;
#pragma omp parallel
    {
        Vec Wt(N), v(N);
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m) {
            const double* x = Z + size_t(m) * ns * N;  // x1 = W1, x2, (Wperp)
            for (int h = 0; h < par_size; ++h) {
                const double* L = Ls[h].data();
                for (int i = 0; i < N; ++i) {  // Wtilde_i = row n+i of L times (x1, x2)
                    double w = 0.0;
                    for (int j = 0; j <= N + i; ++j) w += L[size_t(N + i) * 2 * N + j] * x[j];
                    Wt[i] = w;
                }
                for (int t = 0; t < op_size; ++t) {
                    orc_update_v(Wt.data(), xi, H[h], eta[h], T[t], N, comp_mode, v.data());
                    const double dt = T[t] / N, sdt = std::sqrt(dt);
                    double pay;
                    if (method == 1) {
                        // log-Euler with Z = rho W1 + sqrt(1 - rho^2) Wperp (updateZ/updateS)
                        const double c = std::sqrt(1.0 - rho[h] * rho[h]);
                        double lnS = 0.0;
                        for (int i = 0; i < N; ++i)
                            lnS += std::sqrt(v[i]) * sdt * (rho[h] * x[i] + c * x[2 * N + i]) - 0.5 * v[i] * dt;
                        const double S = std::exp(lnS);
                        pay = S > K[t] ? S - K[t] : 0.0;
                    } else {
                        const double Iv = orc_int_vdt(v.data(), N, dt);
                        const double Is = orc_int_rootvdw(v.data(), x, N, sdt);
                        pay = orc_rt_payoff(Iv, Is, rho[h], K[t]);
                    }
                    payoff[m * out + long(h) * op_size + t] = pay;
                }
            }
        }
    }
    return 0;
}

void orc_payoff_mc_samples(double xi, const double* H, const double* eta, const double* rho,
                           int par_size, const double* T, const double* K, int op_size,
                           int N, long M, const double* Z, int comp_mode, double* payoff) {
    Params par{H, eta, rho, T, K, par_size, op_size};
    const std::vector<Vec> g = gamma_tables(par, N);
    const long out = par.size();
#pragma omp parallel
    {
        PathState st(N);
        double S = 0.0;
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m) {
            const double* W1 = Z + 3 * m * N;
            const double* W1p = W1 + N;
            const double* Wp = W1 + 2 * N;
            /* rBergLib/rBergomiMT.cpp:238-267 */
            for (long I = 0; I < out; ++I) {
                bool update = par.HTrig(I);
                if (update) wtilde_impl(W1, W1p, g[I / op_size].data(), par.h(I), N, nullptr, st.Wt.data());
                update = update || par.TTrig(I);
                if (update) {
                    st.dt = par.t(I) / N;
                    st.sdt = std::sqrt(st.dt);
                }
                update = update || par.etaTrig(I);
                if (update) v_impl(st.Wt.data(), xi, par.h(I), par.e(I), par.t(I), N, comp_mode, st.v.data());
                update = update || par.rhoTrig(I);
                if (update) S = s_impl(st.v.data(), W1, Wp, par.r(I), st.dt, N);
                payoff[m * out + I] = posPart(S - par.k(I));
            }
        }
    }
}

void orc_price_stat(const double* payoff, long M, int out_size, double* price, double* stat) {
    std::vector<double> s(out_size, 0.0), s2(out_size, 0.0);
    for (long m = 0; m < M; ++m)
        for (int I = 0; I < out_size; ++I) {
            const double x = payoff[m * out_size + I];
            s[I] += x;
            s2[I] += x * x;
        }
    for (int I = 0; I < out_size; ++I) {
        price[I] = s[I] / double(M);
        const double var = s2[I] / double(M) - price[I] * price[I];
        stat[I] = std::sqrt(var) / std::sqrt(double(M));
    }
}

/* Shared driver of orc_compute_rt / orc_compute_mc (rBergLib/rBergomiMT.cpp:174-289 and
 * :301-414; single thread rBergLib/RBergomi.cpp:218-370). */
static void compute_seeded(bool rt, double xi, const double* H, const double* eta,
                           const double* rho, int par_size, const double* T, const double* K,
                           int op_size, int N, long M, const uint64_t* seed, int seed_size,
                           int num_threads, int conv_mode, double* price, double* iv,
                           double* stat) {
    Params par{H, eta, rho, T, K, par_size, op_size};
    const std::vector<Vec> g = gamma_tables(par, N);
    const long out = par.size();
    if (num_threads < 1) num_threads = 1;
    /* seeding: single thread = RBergomiST::setGen (seed_seq over the whole seed array,
     * generate seed_size words, take the first); several threads = RNG::RNG (generate one
     * word per thread). */
    std::vector<uint64_t> seeds(std::max(num_threads, 1));
    {
        std::seed_seq seeder(seed, seed + seed_size);
        if (num_threads == 1) {
            std::vector<uint64_t> tmp(seed, seed + seed_size);
            seeder.generate(tmp.begin(), tmp.end());
            seeds[0] = tmp.empty() ? 0 : tmp[0];
        } else {
            seeder.generate(seeds.begin(), seeds.end());
        }
    }
    std::vector<double> sum(out, 0.0), sum2(out, 0.0);
    /* thread-private partial sums, combined in thread order after the loop instead of in
     * the arrival order of the reference's critical section (:395-401), so that a fixed
     * thread count gives bit-identical results */
    std::vector<Vec> part(num_threads), part2(num_threads);
#ifdef _OPENMP
    omp_set_dynamic(0);
#endif
#pragma omp parallel num_threads(num_threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::mt19937_64 gen(seeds[tid]);
        std::normal_distribution<double> dist(0.0, 1.0);
        Vec W1(N), W1perp(N), Wperp(N), pay(out);
        Vec ps(out, 0.0), ps2(out, 0.0);
        PathState st(N);
        FFTWork fft(N);
        FFTWork* f = conv_mode == ORC_CONV_FFT ? &fft : nullptr;
        double S = 0.0;
#pragma omp for schedule(static)
        for (long m = 0; m < M; ++m) {
            for (auto& x : W1) x = dist(gen);
            for (auto& x : W1perp) x = dist(gen);
            if (rt) {
                rt_path(par, xi, N, W1.data(), W1perp.data(), ORC_COMP_REF, f, g, st, pay.data(), nullptr);
            } else {
                for (auto& x : Wperp) x = dist(gen);
                for (long I = 0; I < out; ++I) {
                    bool update = par.HTrig(I);
                    if (update) wtilde_impl(W1.data(), W1perp.data(), g[I / op_size].data(), par.h(I), N, f, st.Wt.data());
                    update = update || par.TTrig(I);
                    if (update) {
                        st.dt = par.t(I) / N;
                        st.sdt = std::sqrt(st.dt);
                    }
                    update = update || par.etaTrig(I);
                    if (update) v_impl(st.Wt.data(), xi, par.h(I), par.e(I), par.t(I), N, ORC_COMP_REF, st.v.data());
                    update = update || par.rhoTrig(I);
                    if (update) S = s_impl(st.v.data(), W1.data(), Wperp.data(), par.r(I), st.dt, N);
                    pay[I] = posPart(S - par.k(I));
                }
            }
            for (long I = 0; I < out; ++I) {
                ps[I] += pay[I];
                ps2[I] += pay[I] * pay[I];
            }
        }
        part[tid] = ps;
        part2[tid] = ps2;
    }
    for (int t = 0; t < num_threads; ++t)
        for (long I = 0; I < out && !part[t].empty(); ++I) {
            sum[I] += part[t][I];
            sum2[I] += part2[t][I];
        }
    for (long I = 0; I < out; ++I) {
        price[I] = sum[I] / double(M);
        const double var = sum2[I] / double(M) - price[I] * price[I];
        stat[I] = std::sqrt(var) / std::sqrt(double(M));
        if (iv) iv[I] = orc_iv_call(price[I], 1.0, par.k(I), par.t(I));
    }
}

void orc_compute_rt(double xi, const double* H, const double* eta, const double* rho,
                    int par_size, const double* T, const double* K, int op_size, int N, long M,
                    const uint64_t* seed, int seed_size, int num_threads, int conv_mode,
                    double* price, double* iv, double* stat) {
    compute_seeded(true, xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size,
                   num_threads, conv_mode, price, iv, stat);
}

void orc_compute_mc(double xi, const double* H, const double* eta, const double* rho,
                    int par_size, const double* T, const double* K, int op_size, int N, long M,
                    const uint64_t* seed, int seed_size, int num_threads, int conv_mode,
                    double* price, double* iv, double* stat) {
    compute_seeded(false, xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size,
                   num_threads, conv_mode, price, iv, stat);
}

/* Philox4x32 round function with the published multipliers and Weyl key increments
 * (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 4.3). */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_key(const uint64_t* seed, int seed_size, uint32_t* key) {
    uint64_t h = splitmix64(0x243F6A8885A308D3ULL ^ uint64_t(seed_size));
    for (int i = 0; i < seed_size; ++i) h = splitmix64(h ^ seed[i]);
    key[0] = uint32_t(h);
    key[1] = uint32_t(h >> 32);
}

void orc_normals_philox(const uint64_t* seed, int seed_size, long p0, long np, int nstreams,
                        int N, double* out) {
    uint32_t key[2];
    orc_philox_key(seed, seed_size, key);
    const double two32i = 1.0 / 4294967296.0;
#pragma omp parallel for schedule(static)
    for (long pp = 0; pp < np; ++pp) {
        const uint64_t p = uint64_t(p0 + pp);
        for (int s = 0; s < nstreams; ++s) {
            double* z = out + (pp * nstreams + s) * long(N);
            for (int q = 0; q < (N + 3) / 4; ++q) {
                const uint32_t ctr[4] = {uint32_t(q), uint32_t(p), uint32_t(p >> 32), uint32_t(s)};
                uint32_t x[4];
                orc_philox4x32_10(ctr, key, x);
                double zz[4];
                for (int j = 0; j < 2; ++j) {
                    const double u0 = (double(x[2 * j]) + 0.5) * two32i;
                    const double u1 = (double(x[2 * j + 1]) + 0.5) * two32i;
                    const double r = std::sqrt(-2.0 * std::log(u0));
                    zz[2 * j] = r * std::cos(2.0 * M_PI * u1);
                    zz[2 * j + 1] = r * std::sin(2.0 * M_PI * u1);
                }
                for (int j = 0; j < 4 && 4 * q + j < N; ++j) z[4 * q + j] = zz[j];
            }
        }
    }
}

/* Brownian-bridge (hierarchical) representation, rBergomi/aux.h:105-125: increment i of the
 * 2^m cells is 2^(-m/2) Z_0 plus, for every level n = 1..m, +-2^((n-m-1)/2) times the
 * coefficient of the level-n cell containing i (+ in its left half, - in its right half). */
void orc_bridge(const double* Z, int m, double* dW) {
    const long M = 1L << m;
    for (long i = 0; i < M; ++i) {
        double v = std::pow(2.0, -0.5 * m) * Z[0];
        for (int n = 1; n <= m; ++n) {
            const long cell = i >> (m - n);                 // floor(i 2^(n-m))
            const bool left = (cell % 2) == 0;
            const long index = (1L << (n - 1)) + cell / 2;  // coefficient of that level-n cell
            v += std::pow(2.0, 0.5 * (n - m - 1)) * (left ? 1.0 : -1.0) * Z[index];
        }
        dW[i] = v;
    }
}

}  // extern "C"
