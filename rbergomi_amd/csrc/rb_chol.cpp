// rb_chol.cpp -- exact joint Gaussian sampling of (W1, Wtilde) by Cholesky factorisation (host).
//
// Replaces the covariance and factorisation of RfBm (PremiaRBergomi/RfBm.cpp:18-80).  On the grid
// t_i = (i+1)/n of [0, 1], W1_j are the unit-variance increments sqrt(n)(B_{(j+1)/n} - B_{j/n}) and
// Wtilde_i = sqrt(2H) int_0^{t_i} (t_i - s)^(H-1/2) dB_s.  Their covariance:
//   Cov(W1_j, W1_k)       = delta_jk                                              (cBm, :39-41)
//   Cov(W1_j, Wtilde_i)   = 2 sqrt(2Hn) ((t - u)^(H+1/2) - (t - v)^(H+1/2)) / (1 + 2H),
//                           u = j/n, v = (j+1)/n, t = t_i, for j <= i, else 0       (cMixed, :49-58)
//   Cov(Wtilde_i, Wtilde_k) = t^(2H) G(s/t), t = min, s = max of t_i, t_k          (cfBm, :43-47)
//   G(x) = 2H (x^-g / (1 - g) + g x^-(1+g) 2F1(1, 1+g; 3-g; 1/x) / ((1-g)(2-g))), g = 1/2 - H (:31-37)
// Because the W1 block is the identity, the factor of the 2n x 2n covariance is
//   L = [[I, 0], [A21, chol(A22 - A21 A21^T)]],
// so only the n x n Schur complement is factorised (the reference factorises all 2n x 2n).
// 2F1 is evaluated with its Gauss series for z <= 1/2 and the linear transformation to 1 - z
// (Abramowitz & Stegun 15.3.6) above, which needs c - a - b = 2H non-integer: H in (0, 1/2).
// (H = 1/2 is Brownian motion, G = 1, and Wtilde is then a function of W1: no factor exists.)
#include <cmath>
#include <string>
#include <vector>

#include "rb_engine.h"

namespace rb {

namespace {

// Gauss series sum_k (a)_k (b)_k / ((c)_k k!) z^k, |z| <= 1/2
double gauss_series(double a, double b, double c, double z) {
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 400; ++k) {
        term *= (a + k) * (b + k) / ((c + k) * (k + 1.0)) * z;
        sum += term;
        if (std::fabs(term) <= 1e-17 * std::fabs(sum)) break;
    }
    return sum;
}

}  // namespace

double hyp2f1(double a, double b, double c, double z) {
    if (z <= 0.5) return gauss_series(a, b, c, z);
    const double w = 1.0 - z, s = c - a - b;
    const double A = std::tgamma(c) * std::tgamma(s) / (std::tgamma(c - a) * std::tgamma(c - b));
    const double B = std::tgamma(c) * std::tgamma(-s) / (std::tgamma(a) * std::tgamma(b));
    return A * gauss_series(a, b, 1.0 - s, w) + B * std::pow(w, s) * gauss_series(c - a, c - b, 1.0 + s, w);
}

static double G_fun(double H, double x) {
    const double g = 0.5 - H;
    if (g == 0.0) return 1.0;
    return 2.0 * H *
           (std::pow(x, -g) / (1.0 - g) +
            g * std::pow(x, -(1.0 + g)) * hyp2f1(1.0, 1.0 + g, 3.0 - g, 1.0 / x) / ((1.0 - g) * (2.0 - g)));
}

void chol_covariance(double H, int n, double* A21, double* A22) {
    const double c = 2.0 * std::sqrt(2.0 * H * n) / (1.0 + 2.0 * H);
    for (int i = 0; i < n; ++i) {
        const double t = double(i + 1) / n;
        for (int j = 0; j < n; ++j) {
            double m = 0.0;
            if (j <= i) {
                const double u = double(j) / n, v = double(j + 1) / n;
                m = c * (std::pow(t - u, 0.5 + H) - std::pow(t - v, 0.5 + H));
            }
            A21[size_t(i) * n + j] = m;
        }
        for (int k = 0; k <= i; ++k) {
            const double tk = double(k + 1) / n;          // tk <= t
            const double cv = std::pow(tk, 2.0 * H) * G_fun(H, t / tk);
            A22[size_t(i) * n + k] = cv;
            A22[size_t(k) * n + i] = cv;
        }
    }
}

int chol_factor(double H, int n, std::vector<double>& L, std::string& err) {
    if (!(H > 0.0 && H < 0.5)) {
        // at H = 1/2, Wtilde = sum W1 / sqrt(n) is a function of W1: the Schur complement is 0
        err = "the Cholesky scheme needs H in (0, 1/2)";
        return ST_ERR_ARG;
    }
    std::vector<double> A21(size_t(n) * n), S(size_t(n) * n);
    chol_covariance(H, n, A21.data(), S.data());
    // Schur complement S = A22 - A21 A21^T (lower triangle; A21 is lower triangular)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k <= i; ++k) {
            double d = 0.0;
            const double* ri = &A21[size_t(i) * n];
            const double* rk = &A21[size_t(k) * n];
            for (int j = 0; j <= k; ++j) d += ri[j] * rk[j];
            S[size_t(i) * n + k] -= d;
        }
    // in-place Cholesky of the lower triangle (row-oriented)
    for (int k = 0; k < n; ++k) {
        double* rk = &S[size_t(k) * n];
        double d = rk[k];
        for (int j = 0; j < k; ++j) d -= rk[j] * rk[j];
        if (!(d > 0.0)) {
            err = "the joint covariance of (W1, Wtilde) is not positive definite at step " +
                  std::to_string(k) + " (H = " + std::to_string(H) + ", N = " + std::to_string(n) + ")";
            return ST_ERR_ARG;
        }
        rk[k] = std::sqrt(d);
        for (int i = k + 1; i < n; ++i) {
            double* ri = &S[size_t(i) * n];
            double x = ri[k];
            for (int j = 0; j < k; ++j) x -= ri[j] * rk[j];
            ri[k] = x / rk[k];
        }
    }
    // L row i: [A21 row i | L22 row i], zeros above the diagonals
    L.assign(size_t(n) * 2 * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            L[size_t(i) * 2 * n + j] = A21[size_t(i) * n + j];
            L[size_t(i) * 2 * n + n + j] = S[size_t(i) * n + j];
        }
    return ST_OK;
}

}  // namespace rb
