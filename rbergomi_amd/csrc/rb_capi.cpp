// rb_capi.cpp -- exported entry points of librbergomi_mi355x.so.
//
// Legacy C ABI (rBergLib/cinterface.cpp:10-131), the C++ interface
// (rBergLib/interface.cpp:11-52), and the status-returning v2 / plan / probe API.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/interface.h"
#include "../../include/rbergomi_mi355x.h"
#include "rb_engine.h"
#include "rb_kernels.h"

using namespace rb;

struct rb_plan {
    Plan plan;
};

namespace {

Request make_request(double xi, const double* H, const double* eta, const double* rho,
                     int par_size, const double* T, const double* K, int op_size, int N, long M,
                     const uint64_t* seed, int seed_size, const rb_config* cfg) {
    Request r;
    r.xi = xi;
    if (par_size > 0) {
        r.H.assign(H, H + par_size);
        r.eta.assign(eta, eta + par_size);
        r.rho.assign(rho, rho + par_size);
    }
    if (op_size > 0) {
        r.T.assign(T, T + op_size);
        r.K.assign(K, K + op_size);
    }
    r.N = N;
    r.M = M;
    if (seed && seed_size > 0) r.seed.assign(seed, seed + seed_size);
    if (cfg) {
        r.precision = cfg->precision;
        r.comp = cfg->compensator;
        r.method = cfg->method;
        r.scheme = cfg->scheme;
    }
    return r;
}

int fail(int rc, const std::string& msg) {
    set_last_error(msg);
    return rc;
}

// unordered rows: one (xi, H, eta, rho, T, K) per output
Request make_rows_request(const double* xi, const double* H, const double* eta, const double* rho,
                          const double* T, const double* K, int rows, int N, long M,
                          const uint64_t* seed, int seed_size, const rb_config* cfg) {
    Request r = make_request(0.0, H, eta, rho, rows, T, K, rows, N, M, seed, seed_size, cfg);
    r.rows = true;
    r.xi_rows.assign(xi, xi + rows);
    return r;
}

int check_rows(const double* xi, const double* H, const double* eta, const double* rho,
               const double* T, const double* K, int rows) {
    if (rows < 1) return fail(RB_ERR_ARG, "rows must be positive");
    if (!xi || !H || !eta || !rho || !T || !K) return fail(RB_ERR_ARG, "null input array");
    return RB_OK;
}

// Legacy path: the reference checks out_size and num_threads and exits (cinterface.cpp:15-20,
// :38-42); device failures have no channel in that ABI, so they exit too (status 40).
void legacy_price(int method, double xi, double* H, double* eta, double* rho, int par_size,
                  double* T, double* K, int op_size, int N, long M, uint64_t* seed,
                  int seed_size, int num_threads, double* price, double* iv, double* stat,
                  double* H_tot, double* eta_tot, double* rho_tot, double* T_tot,
                  double* K_tot, int out_size) {
    if (out_size != par_size * op_size) {
        std::fprintf(stderr,
                     "Output arrays have wrong size:\npar_size * op_size = %d != out_size.\n",
                     par_size * op_size);
        std::exit(17);
    }
    if (num_threads < 1) {
        std::fprintf(stderr,
                     "Number of threads has to be a positive integer,\nobtained num_threads = %d.\n",
                     num_threads);
        std::exit(21);
    }
    rb_config cfg{RB_PREC_FP64, RB_COMP_REF, method, 0};
    std::string err;
    Request r = make_request(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, &cfg);
    int rc = validate(r, err);
    if (rc == ST_ERR_ARG) {
        // the reference does not validate values and returns NaN (e.g. H = 0 divides by
        // sqrt(2H) in updateWtilde); keep that contract and say why on stderr
        std::fprintf(stderr, "rbergomi_mi355x: %s; results set to NaN\n", err.c_str());
        for (int I = 0; I < out_size; ++I) price[I] = iv[I] = stat[I] = std::nan("");
    } else {
        Plan p;
        if (!rc) rc = p.init(r, nullptr, false, err);
        if (!rc) rc = p.run(err);
        if (!rc) rc = p.results(price, iv, stat, err);
        if (rc) {
            std::fprintf(stderr, "rbergomi_mi355x: %s\n", err.c_str());
            std::exit(40);
        }
    }
    // Result2Array (cinterface.cpp:88-100); computeIV_tensor passes rho_tot/eta_tot swapped (:45)
    double* e_out = method == RB_METHOD_MC ? rho_tot : eta_tot;
    double* r_out = method == RB_METHOD_MC ? eta_tot : rho_tot;
    for (int I = 0; I < out_size; ++I) {
        const int h = I / op_size, t = I % op_size;
        H_tot[I] = H[h];
        r_out[I] = rho[h];
        e_out[I] = eta[h];
        T_tot[I] = T[t];
        K_tot[I] = K[t];
    }
}

ResultSimple cpp_price(int method, double xi, const Vector& H, const Vector& eta,
                       const Vector& rho, const Vector& T, const Vector& K, int N, long M,
                       int numThreads, const std::vector<uint64_t>& seed) {
    rb_config cfg{RB_PREC_FP64, RB_COMP_REF, method, 0};
    const int par = int(H.size()), op = int(T.size()), out = par * op;
    ResultSimple res;
    res.price.assign(out, 0.0);
    res.iv.assign(out, 0.0);
    res.stat.assign(out, 0.0);
    const int rc = rb_price_v2(xi, H.data(), eta.data(), rho.data(), par, T.data(), K.data(), op, N,
                               M, seed.data(), int(seed.size()), &cfg, res.price.data(),
                               res.iv.data(), res.stat.data(), out);
    if (rc != RB_OK) throw std::runtime_error(std::string("rbergomi_mi355x: ") + rb_last_error());
    // Result2ResultSimple, rBergLib/interface.cpp:37-52
    for (int I = 0; I < out; ++I) {
        res.H.push_back(H[I / op]);
        res.rho.push_back(rho[I / op]);
        res.eta.push_back(eta[I / op]);
        res.T.push_back(T[I % op]);
        res.K.push_back(K[I % op]);
    }
    res.xi = xi;
    res.N = N;
    res.M = M;
    res.numThreads = numThreads;
    return res;
}

}  // namespace

extern "C" {

void computeIV_tensor(double xi, double* H, double* eta, double* rho, int par_size, double* T,
                      double* K, int op_size, int N, long M, uint64_t* seed, int seed_size,
                      int num_threads, double* price, double* iv, double* stat, double* H_tot,
                      double* eta_tot, double* rho_tot, double* T_tot, double* K_tot,
                      int out_size) {
    legacy_price(RB_METHOD_MC, xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size,
                 num_threads, price, iv, stat, H_tot, eta_tot, rho_tot, T_tot, K_tot, out_size);
}

void computeIVRT_tensor(double xi, double* H, double* eta, double* rho, int par_size, double* T,
                        double* K, int op_size, int N, long M, uint64_t* seed, int seed_size,
                        int num_threads, double* price, double* iv, double* stat, double* H_tot,
                        double* eta_tot, double* rho_tot, double* T_tot, double* K_tot,
                        int out_size) {
    legacy_price(RB_METHOD_RT, xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size,
                 num_threads, price, iv, stat, H_tot, eta_tot, rho_tot, T_tot, K_tot, out_size);
}

void computeIVRT_tensor_quadrature(double xi, double* H, double* eta, double* rho, int par_size,
                                   double* T, double* K, int op_size, int N, long M, double* Z,
                                   int num_threads, double* payoff) {
    (void)num_threads;
    rb_config cfg{RB_PREC_FP64, RB_COMP_REF, RB_METHOD_RT, 0};
    const int rc = rb_payoff_samples_v2(xi, H, eta, rho, par_size, T, K, op_size, N, M, Z, &cfg,
                                        payoff, nullptr, nullptr);
    if (rc != RB_OK) {
        std::fprintf(stderr, "rbergomi_mi355x: %s\n", rb_last_error());
        std::exit(40);
    }
}

const char* rb_last_error(void) { return last_error(); }

int rb_price_v2(double xi, const double* H, const double* eta, const double* rho, int par_size,
                const double* T, const double* K, int op_size, int N, long M,
                const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                double* iv, double* stat, int out_size) {
    set_last_error("");
    if (out_size != par_size * op_size) return fail(RB_ERR_SIZE, "out_size != par_size*op_size");
    if (par_size < 1 || op_size < 1) return fail(RB_ERR_ARG, "par_size and op_size must be positive");
    if (!H || !eta || !rho || !T || !K || !price || !stat)
        return fail(RB_ERR_ARG, "null input or output array");
    rb_plan* p = nullptr;
    int rc = rb_plan_create(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, cfg,
                            nullptr, &p);
    if (rc) return rc;
    rc = rb_plan_run(p);
    if (!rc) rc = rb_plan_results(p, price, iv, stat);
    rb_plan_destroy(p);
    return rc;
}

static int price_sobol_impl(double xi, const double* H, const double* eta, const double* rho, int par_size,
                            const double* T, const double* K, int op_size, int N, long M,
                            const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                            double* iv, double* stat, int out_size, int gauss) {
    set_last_error("");
    if (out_size != par_size * op_size) return fail(RB_ERR_SIZE, "out_size != par_size*op_size");
    if (par_size < 1 || op_size < 1) return fail(RB_ERR_ARG, "par_size and op_size must be positive");
    if (!H || !eta || !rho || !T || !K || !price || !stat) return fail(RB_ERR_ARG, "null array");
    Request r = make_request(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, cfg);
    r.gauss = gauss;
    std::string err;
    Plan p;
    int rc = p.init(r, nullptr, false, err);
    if (!rc) rc = p.run(err);
    if (!rc) rc = p.results(price, iv, stat, err);
    if (rc) return fail(rc, err);
    return RB_OK;
}

int rb_price_sobol(double xi, const double* H, const double* eta, const double* rho, int par_size,
                   const double* T, const double* K, int op_size, int N, long M,
                   const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                   double* iv, double* stat, int out_size) {
    return price_sobol_impl(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, cfg, price, iv,
                            stat, out_size, 1);
}

int rb_price_sobol_bridge(double xi, const double* H, const double* eta, const double* rho, int par_size,
                          const double* T, const double* K, int op_size, int N, long M,
                          const uint64_t* seed, int seed_size, const rb_config* cfg, double* price,
                          double* iv, double* stat, int out_size) {
    return price_sobol_impl(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, cfg, price, iv,
                            stat, out_size, 2);
}

int rb_probe_sobol(int dims, long i0, long n, double* u) {
    set_last_error("");
    if (dims < 1 || n < 1 || i0 < 0 || !u) return fail(RB_ERR_ARG, "bad arguments");
    std::vector<uint32_t> V;
    sobol_directions(dims, V);
    std::vector<double> out(size_t(n) * dims);
    for (long p = 0; p < n; ++p)
        for (int d = 0; d < dims; ++d) {
            const uint64_t i = uint64_t(i0 + p);
            uint32_t x = 0;
            for (int k = 0; k < 32; ++k)
                if (i >> k & 1) x ^= V[size_t(d) * 32 + k];
            out[size_t(p) * dims + d] = double(x) * 2.3283064365386963e-10;
        }
    std::copy(out.begin(), out.end(), u);
    return RB_OK;
}

int rb_probe_norm_inv(const double* u, long n, double* z) {
    set_last_error("");
    if (!u || !z || n < 1) return fail(RB_ERR_ARG, "bad arguments");
    double *du = nullptr, *dz = nullptr;
    hipError_t e = hipMalloc(&du, n * sizeof(double));
    if (!e) e = hipMalloc(&dz, n * sizeof(double));
    if (!e) e = hipMemcpy(du, u, n * sizeof(double), hipMemcpyHostToDevice);
    if (!e) e = launch_norm_inv_probe(du, n, dz, nullptr);
    if (!e) e = hipMemcpy(z, dz, n * sizeof(double), hipMemcpyDeviceToHost);
    if (du) (void)hipFree(du);
    if (dz) (void)hipFree(dz);
    return e ? fail(RB_ERR_DEVICE, hipGetErrorString(e)) : RB_OK;
}

int rb_iv_reference(const double* price, const double* K, const double* T, int n, double* iv) {
    set_last_error("");
    if (!price || !K || !T || !iv || n < 1) return fail(RB_ERR_ARG, "bad arguments");
    const size_t b = size_t(n) * sizeof(double);
    double *dp = nullptr, *dk = nullptr, *dt = nullptr, *dv = nullptr;
    hipError_t e = hipMalloc(&dp, b);
    if (!e) e = hipMalloc(&dk, b);
    if (!e) e = hipMalloc(&dt, b);
    if (!e) e = hipMalloc(&dv, b);
    if (!e) e = hipMemcpy(dp, price, b, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(dk, K, b, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(dt, T, b, hipMemcpyHostToDevice);
    if (!e) e = launch_iv_ref_f64(dp, dk, dt, n, dv, nullptr);
    if (!e) e = hipMemcpy(iv, dv, b, hipMemcpyDeviceToHost);
    for (double* x : {dp, dk, dt, dv})
        if (x) (void)hipFree(x);
    return e ? fail(RB_ERR_DEVICE, hipGetErrorString(e)) : RB_OK;
}

int rb_payoff_samples_v2(double xi, const double* H, const double* eta, const double* rho,
                         int par_size, const double* T, const double* K, int op_size, int N,
                         long M, const double* Z, const rb_config* cfg, double* payoff,
                         double* price, double* stat) {
    set_last_error("");
    if (par_size < 1 || op_size < 1) return fail(RB_ERR_ARG, "par_size and op_size must be positive");
    if (!H || !eta || !rho || !T || !K || !Z || !payoff) return fail(RB_ERR_ARG, "null array");
    Request r = make_request(xi, H, eta, rho, par_size, T, K, op_size, N, M, nullptr, 0, cfg);
    std::string err;
    Plan p;
    int rc = p.init(r, nullptr, true, err);
    if (!rc) rc = p.run_samples(Z, payoff, err);
    if (!rc && (price || stat)) rc = p.results(price, nullptr, stat, err);
    if (rc) return fail(rc, err);
    return RB_OK;
}

int rb_plan_create(double xi, const double* H, const double* eta, const double* rho,
                   int par_size, const double* T, const double* K, int op_size, int N, long M,
                   const uint64_t* seed, int seed_size, const rb_config* cfg, void* stream,
                   rb_plan** out) {
    set_last_error("");
    if (!out) return fail(RB_ERR_ARG, "null plan pointer");
    *out = nullptr;
    if (par_size < 1 || op_size < 1) return fail(RB_ERR_ARG, "par_size and op_size must be positive");
    if (!H || !eta || !rho || !T || !K) return fail(RB_ERR_ARG, "null input array");
    std::unique_ptr<rb_plan> p(new rb_plan());
    Request r = make_request(xi, H, eta, rho, par_size, T, K, op_size, N, M, seed, seed_size, cfg);
    std::string err;
    const int rc = p->plan.init(r, static_cast<hipStream_t>(stream), false, err);
    if (rc) return fail(rc, err);
    *out = p.release();
    return RB_OK;
}

int rb_plan_create_rows(const double* xi, const double* H, const double* eta, const double* rho,
                        const double* T, const double* K, int rows, int N, long M,
                        const uint64_t* seed, int seed_size, const rb_config* cfg, void* stream,
                        rb_plan** out) {
    set_last_error("");
    if (!out) return fail(RB_ERR_ARG, "null plan pointer");
    *out = nullptr;
    int rc = check_rows(xi, H, eta, rho, T, K, rows);
    if (rc) return rc;
    std::unique_ptr<rb_plan> p(new rb_plan());
    Request r = make_rows_request(xi, H, eta, rho, T, K, rows, N, M, seed, seed_size, cfg);
    std::string err;
    rc = p->plan.init(r, static_cast<hipStream_t>(stream), false, err);
    if (rc) return fail(rc, err);
    *out = p.release();
    return RB_OK;
}

int rb_price_rows(const double* xi, const double* H, const double* eta, const double* rho,
                  const double* T, const double* K, int rows, int N, long M, const uint64_t* seed,
                  int seed_size, const rb_config* cfg, double* price, double* iv, double* stat) {
    set_last_error("");
    if (!price || !stat) return fail(RB_ERR_ARG, "null output array");
    rb_plan* p = nullptr;
    int rc = rb_plan_create_rows(xi, H, eta, rho, T, K, rows, N, M, seed, seed_size, cfg, nullptr, &p);
    if (rc) return rc;
    rc = rb_plan_run(p);
    if (!rc) rc = rb_plan_results(p, price, iv, stat);
    rb_plan_destroy(p);
    return rc;
}

int rb_price_rows_turbo(const double* xi, const double* H, const double* eta, const double* rho,
                        const double* T, const double* K, int rows, int N, long M,
                        const uint64_t* seed, int seed_size, long m_control, double q_max,
                        double* price, double* iv, double* stat, double* alpha, double* Q,
                        long* exceptions) {
    set_last_error("");
    int rc = check_rows(xi, H, eta, rho, T, K, rows);
    if (rc) return rc;
    if (!price || !stat) return fail(RB_ERR_ARG, "null output array");
    Request r = make_rows_request(xi, H, eta, rho, T, K, rows, N, M, seed, seed_size, nullptr);
    r.turbo = true;
    r.m_control = m_control;
    r.q_max = q_max;
    std::string err;
    Plan p;
    rc = p.init(r, nullptr, false, err);
    if (!rc) rc = p.run(err);
    if (!rc) rc = p.results(price, iv, stat, err);
    if (!rc && (alpha || Q || exceptions)) rc = p.turbo_results(alpha, Q, exceptions, err);
    if (rc) return fail(rc, err);
    return RB_OK;
}

int rb_payoff_rows_samples(const double* xi, const double* H, const double* eta, const double* rho,
                           const double* T, const double* K, int rows, int N, long M,
                           const double* Z, const rb_config* cfg, double* payoff, double* price,
                           double* stat) {
    set_last_error("");
    int rc = check_rows(xi, H, eta, rho, T, K, rows);
    if (rc) return rc;
    if (!Z || !payoff) return fail(RB_ERR_ARG, "null array");
    Request r = make_rows_request(xi, H, eta, rho, T, K, rows, N, M, nullptr, 0, cfg);
    std::string err;
    Plan p;
    rc = p.init(r, nullptr, true, err);
    if (!rc) rc = p.run_samples(Z, payoff, err);
    if (!rc && (price || stat)) rc = p.results(price, nullptr, stat, err);
    if (rc) return fail(rc, err);
    return RB_OK;
}

int rb_plan_run(rb_plan* plan) {
    if (!plan) return fail(RB_ERR_ARG, "null plan");
    std::string err;
    const int rc = plan->plan.run(err);
    return rc ? fail(rc, err) : RB_OK;
}

int rb_plan_results(rb_plan* plan, double* price, double* iv, double* stat) {
    if (!plan) return fail(RB_ERR_ARG, "null plan");
    std::string err;
    const int rc = plan->plan.results(price, iv, stat, err);
    return rc ? fail(rc, err) : RB_OK;
}

int rb_plan_set_profiling(rb_plan* plan, int enable) {
    if (!plan) return fail(RB_ERR_ARG, "null plan");
    plan->plan.profile = enable != 0;
    return RB_OK;
}

int rb_plan_set_graph(rb_plan* plan, int enable) {
    if (!plan) return fail(RB_ERR_ARG, "null plan");
    plan->plan.use_graph = enable != 0;
    return RB_OK;
}

int rb_plan_kernel_times(rb_plan* plan, char* names, double* total_ms, int* launches, int max,
                         int* count) {
    if (!plan || !count) return fail(RB_ERR_ARG, "null argument");
    int n = 0;
    for (const KernelTime& t : plan->plan.times) {
        if (t.launches == 0) continue;
        if (n >= max) break;
        if (names) {
            std::strncpy(names + 32 * n, t.name.c_str(), 31);
            names[32 * n + 31] = 0;
        }
        if (total_ms) total_ms[n] = t.total_ms;
        if (launches) launches[n] = t.launches;
        ++n;
    }
    *count = n;
    return RB_OK;
}

int rb_plan_info(rb_plan* plan, int* nunits, int* ngroups, int* noptions, long* chunk) {
    if (!plan) return fail(RB_ERR_ARG, "null plan");
    if (nunits) *nunits = int(plan->plan.lay.units.size());
    if (ngroups) *ngroups = int(plan->plan.lay.groups.size());
    if (noptions) *noptions = int(plan->plan.lay.options.size());
    if (chunk) *chunk = plan->plan.chunk;
    return RB_OK;
}

void rb_plan_destroy(rb_plan* plan) { delete plan; }

int rb_layout(const double* H, const double* eta, const double* rho, int par_size,
              const double* T, const double* K, int op_size, int N, int method, int* nunits,
              int* ngroups, int* group_of_option, int* unit_of_group) {
    rb_config cfg{RB_PREC_FP64, RB_COMP_REF, method, 0};
    Request r = make_request(0.04, H, eta, rho, par_size, T, K, op_size, N, 1, nullptr, 0, &cfg);
    std::string err;
    const int rc = validate(r, err);
    if (rc) return fail(rc, err);
    Layout L = make_layout(r);
    if (nunits) *nunits = int(L.units.size());
    if (ngroups) *ngroups = int(L.groups.size());
    if (group_of_option)
        for (size_t I = 0; I < L.options.size(); ++I) group_of_option[I] = L.options[I].group;
    if (unit_of_group)
        for (size_t g = 0; g < L.groups.size(); ++g) unit_of_group[g] = L.groups[g].unit;
    return RB_OK;
}

#define PROBE_TRY(expr)                                                            \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            rc = fail(RB_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
            goto done;                                                             \
        }                                                                          \
    } while (0)

int rb_probe_chol(double H, int N, double* A21, double* A22, double* L) {
    set_last_error("");
    if (N < 1 || N > 2048) return fail(RB_ERR_ARG, "N in [1, 2048]");
    std::vector<double> a21(size_t(N) * N), a22(size_t(N) * N);
    chol_covariance(H, N, a21.data(), a22.data());
    if (A21) std::copy(a21.begin(), a21.end(), A21);
    if (A22) std::copy(a22.begin(), a22.end(), A22);
    if (L) {
        std::vector<double> f;
        std::string err;
        const int rc = chol_factor(H, N, f, err);
        if (rc) return fail(rc, err);
        std::copy(f.begin(), f.end(), L);
    }
    return RB_OK;
}

int rb_probe_philox(const uint32_t* ctr, const uint32_t* key, long n, uint32_t* out) {
    int rc = RB_OK;
    void *dc = nullptr, *dout = nullptr;
    PhiloxKey k{key[0], key[1]};
    PROBE_TRY(hipMalloc(&dc, size_t(n) * 16));
    PROBE_TRY(hipMalloc(&dout, size_t(n) * 16));
    PROBE_TRY(hipMemcpy(dc, ctr, size_t(n) * 16, hipMemcpyHostToDevice));
    PROBE_TRY(launch_philox_probe(dc, k, n, dout, nullptr));
    PROBE_TRY(hipMemcpy(out, dout, size_t(n) * 16, hipMemcpyDeviceToHost));
done:
    if (dc) (void)hipFree(dc);
    if (dout) (void)hipFree(dout);
    return rc;
}

int rb_probe_mfma_layout(const float* A, const float* B, int mode, float* out) {
    int rc = RB_OK;
    if (mode < 0 || mode > 3) {
        set_last_error("mode must be 0 (FP16), 1 (BF16), 2 (FP32) or 3 (FP64 16x16x4)");
        return RB_ERR_ARG;
    }
    float *dA = nullptr, *dB = nullptr, *dO = nullptr;
    PROBE_TRY(hipMalloc(&dA, 512 * sizeof(float)));
    PROBE_TRY(hipMalloc(&dB, 512 * sizeof(float)));
    PROBE_TRY(hipMalloc(&dO, 1024 * sizeof(float)));
    PROBE_TRY(hipMemcpy(dA, A, 512 * sizeof(float), hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(dB, B, 512 * sizeof(float), hipMemcpyHostToDevice));
    PROBE_TRY(launch_probe_mfma_layout(dA, dB, dO, mode, nullptr));
    PROBE_TRY(hipMemcpy(out, dO, 1024 * sizeof(float), hipMemcpyDeviceToHost));
done:
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dO) (void)hipFree(dO);
    return rc;
}

int rb_probe_copy(long bytes, int reps, double* avg_ms) {
    int rc = RB_OK;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    if (bytes < 16 || bytes % 16 || reps < 1) return fail(RB_ERR_ARG, "bytes must be a positive multiple of 16");
    PROBE_TRY(hipMalloc(&a, size_t(bytes)));
    PROBE_TRY(hipMalloc(&b, size_t(bytes)));
    PROBE_TRY(hipMemset(a, 1, size_t(bytes)));
    PROBE_TRY(hipEventCreate(&e0));
    PROBE_TRY(hipEventCreate(&e1));
    PROBE_TRY(launch_probe_copy(a, b, bytes, nullptr));
    PROBE_TRY(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) PROBE_TRY(launch_probe_copy(a, b, bytes, nullptr));
    PROBE_TRY(hipEventRecord(e1, nullptr));
    PROBE_TRY(hipEventSynchronize(e1));
    PROBE_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (avg_ms) *avg_ms = double(ms) / reps;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int rb_probe_issue(double* ns_per_inst, int* simds) {
    int rc = RB_OK;
    float* d = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipDeviceProp_t prop;
    int dev = 0;
    if (!ns_per_inst) return fail(RB_ERR_ARG, "ns_per_inst is null");
    PROBE_TRY(hipGetDevice(&dev));
    PROBE_TRY(hipGetDeviceProperties(&prop, dev));
    if (simds) *simds = 4 * prop.multiProcessorCount;
    PROBE_TRY(hipMalloc(&d, size_t(2048) * 256 * sizeof(float)));
    PROBE_TRY(hipEventCreate(&e0));
    PROBE_TRY(hipEventCreate(&e1));
    for (int op = 0; op < 4; ++op) {
        const int reps = 5;
        float ms = 0.f;
        PROBE_TRY(launch_probe_issue(op, d, 1u, nullptr));  // warm-up
        PROBE_TRY(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) PROBE_TRY(launch_probe_issue(op, d, 2u + uint32_t(r), nullptr));
        PROBE_TRY(hipEventRecord(e1, nullptr));
        PROBE_TRY(hipEventSynchronize(e1));
        PROBE_TRY(hipEventElapsedTime(&ms, e0, e1));
        ns_per_inst[op] = double(ms) / reps * 1e6 / probe_issue_insts_per_simd(4 * prop.multiProcessorCount);
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d) (void)hipFree(d);
    return rc;
}

int rb_probe_payoff(const double* lnS, const double* sig, const double* K, int n, double* out,
                    const float* z, float* mills, int form) {
    int rc = RB_OK;
    double *d0 = nullptr, *d1 = nullptr, *d2 = nullptr, *d3 = nullptr;
    float *f0 = nullptr, *f1 = nullptr;
    const size_t b = size_t(n) * sizeof(double), fb = size_t(n) * sizeof(float);
    PROBE_TRY(hipMalloc(&d0, b));
    PROBE_TRY(hipMalloc(&d1, b));
    PROBE_TRY(hipMalloc(&d2, b));
    PROBE_TRY(hipMalloc(&d3, b));
    PROBE_TRY(hipMalloc(&f0, fb));
    PROBE_TRY(hipMalloc(&f1, fb));
    PROBE_TRY(hipMemcpy(d0, lnS, b, hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(d1, sig, b, hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(d2, K, b, hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(f0, z, fb, hipMemcpyHostToDevice));
    PROBE_TRY(launch_probe_payoff(d0, d1, d2, n, d3, f1, f0, form, nullptr));
    PROBE_TRY(hipMemcpy(out, d3, b, hipMemcpyDeviceToHost));
    PROBE_TRY(hipMemcpy(mills, f1, fb, hipMemcpyDeviceToHost));
done:
    for (void* p : {(void*)d0, (void*)d1, (void*)d2, (void*)d3, (void*)f0, (void*)f1})
        if (p) (void)hipFree(p);
    return rc;
}

int rb_probe_normals(const uint64_t* seed, int seed_size, long p0, long np, int nstreams, int N,
                     double* out) {
    int rc = RB_OK;
    double* d = nullptr;
    const size_t n = size_t(np) * nstreams * N;
    PROBE_TRY(hipMalloc(&d, n * sizeof(double)));
    PROBE_TRY(launch_normals_f64(philox_key(seed, seed_size), p0, np, nstreams, N, d, nullptr));
    PROBE_TRY(hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost));
done:
    if (d) (void)hipFree(d);
    return rc;
}

int rb_probe_bridge(const double* Z, long rows, int N, double* dW) {
    set_last_error("");
    if (!Z || !dW || rows < 1 || N < 1 || (N & (N - 1)) != 0)
        return fail(RB_ERR_ARG, "rows >= 1 of N = 2^m values required");
    int rc = RB_OK;
    double* d = nullptr;
    const size_t n = size_t(rows) * N;
    PROBE_TRY(hipMalloc(&d, n * sizeof(double)));
    PROBE_TRY(hipMemcpy(d, Z, n * sizeof(double), hipMemcpyHostToDevice));
    PROBE_TRY(launch_bridge_f64(d, rows, N, nullptr));
    PROBE_TRY(hipMemcpy(dW, d, n * sizeof(double), hipMemcpyDeviceToHost));
done:
    if (d) (void)hipFree(d);
    return rc;
}

int rb_probe_wtilde(const double* Z, long M, int N, double H, double* Wtilde) {
    const double eta = 1.0, rho = 0.0, T = 1.0, K = 1.0;
    Request r = make_request(0.04, &H, &eta, &rho, 1, &T, &K, 1, N, M, nullptr, 0, nullptr);
    std::string err;
    int rc = validate(r, err);
    if (rc) return fail(rc, err);
    Layout L = make_layout(r);
    double *dZ = nullptr, *dW = nullptr, *dg = nullptr;
    Unit* du = nullptr;
    PROBE_TRY(hipMalloc(&dZ, size_t(M) * 2 * N * sizeof(double)));
    PROBE_TRY(hipMalloc(&dW, size_t(M) * N * sizeof(double)));
    PROBE_TRY(hipMalloc(&dg, size_t(N) * sizeof(double)));
    PROBE_TRY(hipMalloc(&du, sizeof(Unit)));
    PROBE_TRY(hipMemcpy(dZ, Z, size_t(M) * 2 * N * sizeof(double), hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(dg, L.gamma.data(), size_t(N) * sizeof(double), hipMemcpyHostToDevice));
    PROBE_TRY(hipMemcpy(du, L.units.data(), sizeof(Unit), hipMemcpyHostToDevice));
    {
        int dev = 0;
        hipDeviceProp_t prop;
        PROBE_TRY(hipGetDevice(&dev));
        PROBE_TRY(device_props(dev, prop));
        PROBE_TRY(launch_volterra_conv_f64(dZ, 2L * N, M, N, 1, du, dg, dW, prop.sharedMemPerBlock, nullptr));
    }
    PROBE_TRY(hipMemcpy(Wtilde, dW, size_t(M) * N * sizeof(double), hipMemcpyDeviceToHost));
done:
    if (dZ) (void)hipFree(dZ);
    if (dW) (void)hipFree(dW);
    if (dg) (void)hipFree(dg);
    if (du) (void)hipFree(du);
    return rc;
}

int rb_probe_integrals(double xi, const double* H, const double* eta, const double* rho,
                       int par_size, const double* T, const double* K, int op_size, int N,
                       long M, const double* Z, const rb_config* cfg, double* out) {
    Request r = make_request(xi, H, eta, rho, par_size, T, K, op_size, N, M, nullptr, 0, cfg);
    std::string err;
    Plan p;
    int rc = p.init(r, nullptr, true, err);
    if (rc) return fail(rc, err);
    if (p.chunk < M) return fail(RB_ERR_LIMIT, "integrals probe needs all paths in one chunk");
    const int nopt = int(p.lay.options.size());
    std::vector<double> ints(size_t(p.lay.groups.size()) * M * 2);
    {
        hipStream_t st = p.stream;
        hipError_t e = hipMemcpyAsync(p.d_Z, Z, size_t(M) * p.nstreams * N * sizeof(double),
                                      hipMemcpyHostToDevice, st);
        if (!e) e = p.conv(M, long(p.nstreams) * N);
        if (!e) e = launch_scan_f64(r.method, p.d_Z, long(p.nstreams) * N, p.d_Wt, M, N, p.d_groups,
                                    int(p.lay.groups.size()), p.d_comp, p.d_ints, st);
        if (!e) e = hipMemcpyAsync(ints.data(), p.d_ints, ints.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, st);
        if (!e) e = hipStreamSynchronize(st);
        if (e) return fail(RB_ERR_DEVICE, hipGetErrorString(e));
    }
    for (long m = 0; m < M; ++m)
        for (int I = 0; I < nopt; ++I) {
            const int g = p.lay.options[I].group;
            out[2 * (m * nopt + I) + 0] = ints[(size_t(g) * M + m) * 2 + 0];
            out[2 * (m * nopt + I) + 1] = ints[(size_t(g) * M + m) * 2 + 1];
        }
    return RB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- C++ interface
ResultSimple ComputeIV_singleThread(double xi, Vector H, Vector eta, Vector rho, Vector T, Vector K,
                                    int N, long M, std::vector<uint64_t> seed) {
    return cpp_price(RB_METHOD_MC, xi, H, eta, rho, T, K, N, M, 1, seed);
}

ResultSimple ComputeIV_multiThread(double xi, Vector H, Vector eta, Vector rho, Vector T, Vector K,
                                   int N, long M, int numThreads, std::vector<uint64_t> seed) {
    return cpp_price(RB_METHOD_MC, xi, H, eta, rho, T, K, N, M, numThreads, seed);
}

ResultSimple ComputeIVRT_singleThread(double xi, Vector H, Vector eta, Vector rho, Vector T,
                                      Vector K, int N, long M, std::vector<uint64_t> seed) {
    return cpp_price(RB_METHOD_RT, xi, H, eta, rho, T, K, N, M, 1, seed);
}

ResultSimple ComputeIVRT_multiThread(double xi, Vector H, Vector eta, Vector rho, Vector T,
                                     Vector K, int N, long M, int numThreads,
                                     std::vector<uint64_t> seed) {
    return cpp_price(RB_METHOD_RT, xi, H, eta, rho, T, K, N, M, numThreads, seed);
}
