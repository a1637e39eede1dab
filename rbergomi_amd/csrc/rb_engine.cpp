// rb_engine.cpp -- host planner and executor.
//
// Replaces the driver layer of the reference: ComputePriceRTMT / ComputePriceMT
// (rBergLib/rBergomiMT.cpp:174-414), the single-thread class RBergomiST
// (rBergLib/RBergomi.cpp:157-370) and the ParamTot index algebra (rBergLib/ParamTot.h:20-50).
// Instead of walking the flat option index per path and recomputing on "triggers", the
// planner deduplicates the work once (units, groups, options; rb_types.h) and launches each
// stage over all paths of a chunk.
#include "rb_engine.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "rb_kernels.h"

namespace rb {

static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
const char* last_error() { return g_last_error.c_str(); }

const char* kernel_name(int id) {
    static const char* names[K_COUNT] = {"k_normals_f64", "k_volterra_f64", "k_scan_f64",
                                         "k_payoff_f64",  "k_reduce_f64",   "k_iv_f64",
                                         "k_fused"};
    return id >= 0 && id < K_COUNT ? names[id] : "?";
}

hipError_t device_props(int device, hipDeviceProp_t& out) {
    static std::mutex mu;
    static std::map<int, hipDeviceProp_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto it = cache.find(device);
    if (it != cache.end()) {
        out = it->second;
        return hipSuccess;
    }
    const hipError_t e = hipGetDeviceProperties(&out, device);
    if (e == hipSuccess) cache[device] = out;
    return e;
}

void keep_pool_cached(int device) {
    static std::mutex mu;
    static std::set<int> done;
    std::lock_guard<std::mutex> lock(mu);
    if (!done.insert(device).second) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t keep = uint64_t(64) << 30;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
}

int validate(const Request& r, std::string& err) {
    if (r.H.empty() || r.T.empty()) {
        err = "empty parameter or option arrays";
        return ST_ERR_ARG;
    }
    if (r.eta.size() != r.H.size() || r.rho.size() != r.H.size() || r.K.size() != r.T.size()) {
        err = "parameter arrays must be zipped: |H| = |eta| = |rho| and |T| = |K|";
        return ST_ERR_ARG;
    }
    if (r.rows && (r.T.size() != r.H.size() || r.xi_rows.size() != r.H.size())) {
        err = "unordered rows: every parameter array has one entry per row";
        return ST_ERR_ARG;
    }
    if (r.out_size() > 65535 || r.H.size() * (r.rows ? 1 : r.T.size()) > 65535) {
        err = "outputs above the supported maximum of 65535 per call";
        return ST_ERR_LIMIT;
    }
    if (r.scheme != SCHEME_HYBRID && r.scheme != SCHEME_CHOLESKY) {
        err = "unknown scheme";
        return ST_ERR_ARG;
    }
    if (r.N < 1 || r.M < 1) {
        err = "N and M must be positive";
        return ST_ERR_ARG;
    }
    if (r.N > 8192) {
        err = "N above the supported maximum of 8192 steps";
        return ST_ERR_LIMIT;
    }
    for (double h : r.H)
        if (!(h > 0.0 && h < 1.0)) {
            err = "H must lie in (0, 1)";
            return ST_ERR_ARG;
        }
    for (double t : r.T)
        if (!(t > 0.0)) {
            err = "maturities must be positive";
            return ST_ERR_ARG;
        }
    for (double k : r.K)
        if (!(k > 0.0)) {
            err = "strikes must be positive";
            return ST_ERR_ARG;
        }
    for (double x : r.rho)
        if (!(std::fabs(x) <= 1.0)) {
            err = "rho must lie in [-1, 1]";
            return ST_ERR_ARG;
        }
    if (!(r.xi >= 0.0)) {
        err = "xi must be non-negative";
        return ST_ERR_ARG;
    }
    for (double x : r.xi_rows)
        if (!(x >= 0.0)) {
            err = "xi must be non-negative";
            return ST_ERR_ARG;
        }
    return ST_OK;
}

Layout make_layout(const Request& r) {
    Layout L;
    const int N = r.N;
    L.N = N;
    const long P = long(r.H.size()), O = long(r.T.size());
    // units: distinct H (getGamma per distinct H, rBergLib/rBergomiMT.cpp:318-324)
    std::map<double, int> unit_of;
    std::vector<int> unit_of_par(P);
    for (long h = 0; h < P; ++h) {
        auto it = unit_of.find(r.H[h]);
        if (it == unit_of.end()) {
            const int u = int(L.units.size());
            unit_of[r.H[h]] = u;
            const double H = r.H[h];
            const double s2H = std::sqrt(2.0 * H);
            const double rhoH = s2H / (H + 0.5);
            Unit un;
            un.H = H;
            un.a = rhoH / s2H;
            un.b = std::sqrt(1.0 - rhoH * rhoH) / s2H;
            un.sc = s2H * std::pow(1.0 / N, H);
            L.units.push_back(un);
            // Gamma_0 = 0, Gamma_k = ((k+1)^(a+1) - k^(a+1))/(a+1)  (getGamma :143-149)
            const double alpha = H - 0.5;
            L.gamma.push_back(0.0);
            for (int k = 1; k < N; ++k)
                L.gamma.push_back((std::pow(k + 1.0, alpha + 1.0) - std::pow(double(k), alpha + 1.0)) /
                                  (alpha + 1.0));
            unit_of_par[h] = u;
        } else {
            unit_of_par[h] = it->second;
        }
    }
    // groups: distinct (H, eta, T, xi) for RT; (H, eta, T, xi, rho) for plain MC.  Outputs in
    // order: I = h*|T| + t (tensor form, rBergLib/ParamTot.h:32-33), or I = row (unordered rows,
    // rBergomi/ParamTot.h:57-75, per-row xi as ML_console_ST/ParamTot.h:25-43)
    std::map<std::tuple<int, double, double, double, double>, int> group_of;
    const long nout = r.out_size();
    for (long I = 0; I < nout; ++I) {
        {
            const long h = r.rows ? I : I / O;
            const long t = r.rows ? I : I % O;
            const double xi = r.rows ? r.xi_rows[I] : r.xi;
            const int u = unit_of_par[h];
            const double rho_key = r.method == METHOD_MC ? r.rho[h] : 0.0;
            const auto key = std::make_tuple(u, r.eta[h], r.T[t], xi, rho_key);
            auto it = group_of.find(key);
            int g;
            if (it == group_of.end()) {
                g = int(L.groups.size());
                group_of[key] = g;
                const double H = r.H[h], T = r.T[t], eta = r.eta[h];
                Group G;
                G.unit = u;
                G.pad = 0;
                G.xi = xi;
                G.eta = eta;
                G.TH = std::pow(T, H);        // scaleWtilde :163-167
                G.dt = T / N;
                G.sdt = std::sqrt(G.dt);
                G.rho = r.rho[h];
                G.crho = std::sqrt(1.0 - r.rho[h] * r.rho[h]);
                L.groups.push_back(G);
                // compensator of updateV (:73-80): 0.5 eta^2 ((i-1) dt)^(2H), or (i dt) in exact mode
                L.comp.push_back(0.0);
                for (int i = 1; i < N; ++i) {
                    const double tc = r.comp == COMP_EXACT ? i * G.dt : (i - 1) * G.dt;
                    L.comp.push_back(0.5 * eta * eta * std::pow(tc, 2 * H));
                }
            } else {
                g = it->second;
            }
            Option o;
            o.group = g;
            o.pad = 0;
            o.rho = r.rho[h];
            o.K = r.K[t];
            o.lnK = std::log(r.K[t]);
            o.T = r.T[t];
            L.options.push_back(o);
        }
    }
    return L;
}

// splitmix64 finaliser (Steele, Lea, Flood, OOPSLA'14) folded over the seed words
static uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

PhiloxKey philox_key(const uint64_t* seed, int n) {
    uint64_t h = splitmix64(0x243F6A8885A308D3ULL ^ uint64_t(n));
    for (int i = 0; i < n; ++i) h = splitmix64(h ^ seed[i]);
    PhiloxKey k;
    k.k0 = uint32_t(h);
    k.k1 = uint32_t(h >> 32);
    return k;
}

#define RB_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return e_ == hipErrorOutOfMemory ? ST_ERR_NOMEM : ST_ERR_DEVICE;         \
        }                                                                            \
    } while (0)

// ---------------------------------------------------------------- workspace cache
// Device buffers of finished plans are kept for later plans: a one-shot call (rb_price_v2, the
// legacy symbols) otherwise spends far longer mapping and unmapping its staging buffers (up to
// ~6 GB) than computing with them -- measured 4 to 12 s against 0.36 s of kernels for 256
// small requests.  A request takes the smallest free block of its device that is at least its
// size and at most twice it, else allocates; freed blocks beyond WS_KEEP bytes are returned to
// the device, oldest first.  A block is reused only after the plan that released it has
// synchronized its stream.
namespace {
struct WsBlock {
    void* p;
    size_t bytes;
    int device;
};
std::mutex g_ws_mu;
std::deque<WsBlock> g_ws_free;          // released blocks, oldest first
std::map<void*, WsBlock> g_ws_used;
size_t g_ws_free_bytes = 0;
constexpr size_t WS_KEEP = size_t(48) << 30;
constexpr size_t WS_GRAIN = size_t(2) << 20;

// caller holds g_ws_mu
void ws_evict(size_t keep, int device) {
    for (auto it = g_ws_free.begin(); it != g_ws_free.end() && g_ws_free_bytes > keep;) {
        if (device >= 0 && it->device != device) {
            ++it;
            continue;
        }
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != it->device) (void)hipSetDevice(it->device);
        (void)hipFree(it->p);
        if (cur != it->device) (void)hipSetDevice(cur);
        g_ws_free_bytes -= it->bytes;
        it = g_ws_free.erase(it);
    }
}
}  // namespace

hipError_t ws_alloc(void** p, size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + WS_GRAIN - 1) / WS_GRAIN * WS_GRAIN;
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_ws_mu);
    auto best = g_ws_free.end();
    for (auto it = g_ws_free.begin(); it != g_ws_free.end(); ++it)
        if (it->device == device && it->bytes >= bytes && it->bytes <= 2 * bytes &&
            (best == g_ws_free.end() || it->bytes < best->bytes))
            best = it;
    if (best != g_ws_free.end()) {
        *p = best->p;
        g_ws_used[best->p] = *best;
        g_ws_free_bytes -= best->bytes;
        g_ws_free.erase(best);
        return hipSuccess;
    }
    e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {  // give the device back what is cached, then try again
        (void)hipGetLastError();
        ws_evict(0, device);
        e = hipMalloc(p, bytes);
    }
    if (e == hipSuccess) g_ws_used[*p] = WsBlock{*p, bytes, device};
    return e;
}

void ws_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_ws_mu);
    const auto it = g_ws_used.find(p);
    if (it == g_ws_used.end()) return;
    g_ws_free.push_back(it->second);
    g_ws_free_bytes += it->second.bytes;
    g_ws_used.erase(it);
    ws_evict(WS_KEEP, -1);
}

template <class T>
static hipError_t dalloc(T** p, size_t n, hipStream_t) {
    return ws_alloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
}

template <class T>
static hipError_t upload(T** p, const std::vector<T>& v, hipStream_t st) {
    hipError_t e = dalloc(p, v.size(), st);
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

Plan::~Plan() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    fused_batch.reset();  // stream-ordered frees on this plan's stream
    fused_fft.reset();
    if (stream) (void)hipStreamSynchronize(stream);
    void* bufs[] = {d_units, d_gamma, d_groups, d_comp, d_opts,  d_Z,     d_Wt,   d_ints,
                    d_partial, d_price, d_stat, d_iv, d_payoff, d_tw, d_ghat, d_ints2,
                    d_Qg, d_alpha, d_exc, d_count, d_chol, d_sobol, d_shift};
    for (void* b : bufs) ws_free(b);  // the stream has synchronized above
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

int Plan::init(const Request& r, hipStream_t st, bool quadrature, std::string& err) {
    int rc = validate(r, err);
    if (rc) return rc;
    req = r;
    lay = make_layout(r);
    int ndev = 0;
    RB_TRY(hipGetDeviceCount(&ndev));
    if (ndev < 1) {
        err = "no HIP device visible";
        return ST_ERR_DEVICE;
    }
    RB_TRY(hipGetDevice(&device));
    {
        hipDeviceProp_t prop;
        RB_TRY(device_props(device, prop));
        keep_pool_cached(device);
        lds_limit = prop.sharedMemPerBlock;
        const int kind = volterra_conv_kind(r.N, lds_limit);
        const size_t conv_lds = kind == 0 ? volterra_mfma_lds_bytes(r.N)
                                : kind == 1 ? volterra_mfma_long_lds_bytes(r.N, volterra_mfma_long_segment(r.N, lds_limit))
                                            : volterra_f64_lds_bytes(r.N);
        if (conv_lds > prop.sharedMemPerBlock) {
            err = "N = " + std::to_string(r.N) + " needs more LDS than one workgroup can hold";
            return ST_ERR_LIMIT;
        }
    }
    if (st) {
        stream = st;
    } else {
        RB_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
    }
    const bool split = fused_mode_of(r.precision) >= 0;  // a fused-kernel precision
    if (r.turbo && (r.precision != PREC_FP64 || r.method != METHOD_RT || quadrature)) {
        err = "the turbo-charged estimator runs the Romano-Touzi FP64 path on device Gaussians";
        return ST_ERR_ARG;
    }
    if (r.scheme == SCHEME_CHOLESKY && r.precision != PREC_FP64) {
        err = "the Cholesky scheme runs on the FP64 path";
        return ST_ERR_ARG;
    }
    if (r.gauss >= 1 && (r.turbo || quadrature)) {
        err = "Sobol' points are not combined with caller Gaussians or the turbo-charged estimator";
        return ST_ERR_ARG;
    }
    if (r.gauss == 2 && (r.N & (r.N - 1)) != 0) {
        err = "the Brownian-bridge mapping needs N a power of two (hierarchical2increments)";
        return ST_ERR_ARG;
    }
    if (r.gauss >= 1 && r.M > 4294967294L) {
        err = "Sobol' points: M above 2^32 - 2";
        return ST_ERR_LIMIT;
    }
    if (r.scheme == SCHEME_CHOLESKY && r.N > 2048) {
        err = "the Cholesky scheme covers N <= 2048";
        return ST_ERR_LIMIT;
    }
    const bool fft1 = r.precision == PREC_FP32_FFT;  // the single-launch FFT kernel at every N
    if ((fft1 || (split && r.N > 1024)) && r.N <= 8192 && !quadrature && r.gauss == 0) {
        // long paths on device Gaussians (or any N on request): the single-launch FFT kernel
        fused_fft.reset(new FftFused());
        const int rc = fused_fft->build(r, lay, device, stream, err);
        if (rc) return rc;
        nblk = fused_fft->args.nitems;
        chunk = r.M;
        return ST_OK;
    }
    if (r.precision == PREC_FP32 || ((split || fft1) && (r.N > 1024 || quadrature || r.gauss >= 1))) {
        // staged pipeline with the FP32 FFT convolution (long paths, user Gaussians)
        conv_fft = true;
    } else if (split) {
        if (quadrature) {
            err = "user-supplied Gaussians are priced on the FP64 path only";
            return ST_ERR_ARG;
        }
        fused = true;
        fused_mode = fused_mode_of(r.precision);
        return init_fused(err);
    }
    if (r.precision != PREC_FP64 && !conv_fft) {
        err = "precision not available in this build";
        return ST_ERR_ARG;
    }
    const int N = r.N;
    nstreams = r.method == METHOD_MC ? 3 : 2;
    const int nopt = int(lay.options.size());
    const long nunits = long(lay.units.size()), ngroups = long(lay.groups.size());
    // chunk the paths so that the staged buffers stay within ~6 GiB
    if (r.turbo) {
        if (r.precision != PREC_FP64 || r.method != METHOD_RT || quadrature) {
            err = "the turbo-charged estimator runs the Romano-Touzi FP64 path on device Gaussians";
            return ST_ERR_ARG;
        }
        if (r.m_control < 2 || !(r.q_max > 0.0)) {
            err = "turbo-charged estimator: m_control >= 2 and q_max > 0 required";
            return ST_ERR_ARG;
        }
    }
    const double per_path = 8.0 * N * (nstreams + nunits) + 16.0 * ngroups * (r.turbo ? 2 : 1) +
                            (quadrature ? 8.0 * nopt : 0.0);
    long ch = long(6.0e9 / per_path);
    // RBERGOMI_MAX_CHUNK caps the paths per chunk (used to test chunk boundaries)
    if (const char* env = std::getenv("RBERGOMI_MAX_CHUNK")) {
        const long cap = std::atol(env);
        if (cap > 0) ch = std::min(ch, cap);
    }
    ch = std::max<long>(256, (ch / 256) * 256);
    const long Mr = ((r.M + 255) / 256) * 256;
    chunk = std::min(ch, Mr);
    if (r.turbo) {
        // the pilot paths are priced as one chunk (two-pass covariances over all of them)
        const long mc = ((r.m_control + 255) / 256) * 256;
        if (mc > ch) {
            err = "m_control above the paths one chunk holds (" + std::to_string(ch) + ")";
            return ST_ERR_LIMIT;
        }
        chunk = std::max(chunk, mc);
    }
    nblk = (r.M + 255) / 256;
    RB_TRY(upload(&d_units, lay.units, stream));
    RB_TRY(upload(&d_gamma, lay.gamma, stream));
    RB_TRY(upload(&d_groups, lay.groups, stream));
    RB_TRY(upload(&d_comp, lay.comp, stream));
    RB_TRY(upload(&d_opts, lay.options, stream));
    if (conv_fft) {
        rc = init_fft(err);
        if (rc) return rc;
    }
    if (r.gauss >= 1) {
        std::vector<uint32_t> V;
        sobol_directions(nstreams * N, V);
        RB_TRY(upload(&d_sobol, V, stream));
        if (!r.seed.empty()) {
            // digital shift: Philox block (0, dimension, 0, stream 7) of the seed's key
            const PhiloxKey key = philox_key(r.seed.data(), int(r.seed.size()));
            std::vector<uint32_t> sh(size_t(nstreams) * N);
            for (size_t d = 0; d < sh.size(); ++d) sh[d] = philox_block(key, uint64_t(d), 7u, 0u).x;
            RB_TRY(upload(&d_shift, sh, stream));
        }
    }
    if (r.scheme == SCHEME_CHOLESKY) {
        // host factor per unit (distinct H), then one upload
        std::vector<double> all, L;
        for (const Unit& un : lay.units) {
            rc = chol_factor(un.H, N, L, err);
            if (rc) return rc;
            all.insert(all.end(), L.begin(), L.end());
        }
        RB_TRY(upload(&d_chol, all, stream));
    }
    RB_TRY(dalloc(&d_Z, size_t(chunk) * nstreams * N, stream));
    RB_TRY(dalloc(&d_Wt, size_t(nunits) * chunk * N, stream));
    RB_TRY(dalloc(&d_ints, size_t(ngroups) * chunk * 2, stream));
    RB_TRY(dalloc(&d_partial, size_t(nopt) * nblk * 2, stream));
    RB_TRY(dalloc(&d_price, nopt, stream));
    RB_TRY(dalloc(&d_stat, nopt, stream));
    RB_TRY(dalloc(&d_iv, nopt, stream));
    if (quadrature) RB_TRY(dalloc(&d_payoff, size_t(chunk) * nopt, stream));
    if (r.turbo) {
        RB_TRY(dalloc(&d_ints2, size_t(ngroups) * chunk * 2, stream));
        RB_TRY(dalloc(&d_Qg, size_t(ngroups), stream));
        RB_TRY(dalloc(&d_alpha, size_t(nopt), stream));
        RB_TRY(dalloc(&d_exc, size_t(nopt) * nblk, stream));
        RB_TRY(dalloc(&d_count, size_t(nopt), stream));
    }
    RB_TRY(hipStreamSynchronize(stream));
    return ST_OK;
}

void Plan::begin_timing() {
    pending.clear();
    ev_used = 0;
    times.assign(K_COUNT, KernelTime());
    for (int i = 0; i < K_COUNT; ++i) times[i].name = kernel_name(i);
    if (fused_fft) times[K_FUSED].name = "k_fused_fft";
    // the convolution slot holds whichever kernel computes Wtilde
    if (req.scheme == SCHEME_CHOLESKY)
        times[K_VOLTERRA].name = "k_volterra_chol_f64";
    else if (conv_fft)
        times[K_VOLTERRA].name = "k_volterra_fft";
    else if (volterra_conv_kind(req.N, lds_limit) == 0)
        times[K_VOLTERRA].name = "k_volterra_mfma_f64";
    else if (volterra_conv_kind(req.N, lds_limit) == 1)
        times[K_VOLTERRA].name = "k_volterra_mfma_long_f64";
}

int Plan::timed(int id) {
    if (!profile || capturing) return -1;
    while (ev_pool.size() < ev_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ev_pool.push_back(e);
    }
    hipEvent_t a = ev_pool[ev_used], b = ev_pool[ev_used + 1];
    ev_used += 2;
    pending.push_back({id, {a, b}});
    (void)hipEventRecord(a, stream);
    return int(pending.size()) - 1;
}

void Plan::mark(int slot, hipEvent_t, hipEvent_t) {
    if (slot < 0) return;
    (void)hipEventRecord(pending[slot].second.second, stream);
}

void Plan::end_timing() {
    if (!profile) return;
    for (auto& p : pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.second.second) == hipSuccess &&
            hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            times[p.first].total_ms += ms;
            times[p.first].launches += 1;
        }
    }
    pending.clear();
}

#define RB_LAUNCH(id, call)                  \
    do {                                     \
        const int slot_ = timed(id);         \
        RB_TRY(call);                        \
        mark(slot_, nullptr, nullptr);       \
    } while (0)

// Enqueue one pricing run on the plan's stream (no synchronization).
int Plan::enqueue(std::string& err) {
    if (fused) return fused_batch->launch(stream, err, capturing ? nullptr : this);
    if (fused_fft) return fused_fft->launch(stream, err, capturing ? nullptr : this);
    const int N = req.N;
    const int nopt = int(lay.options.size());
    const PhiloxKey key = philox_key(req.seed.data(), int(req.seed.size()));
    const int ngroups = int(lay.groups.size());
    if (req.turbo) {
        // pilot paths (counters TURBO_PILOT_BASE + m, disjoint from the main paths): Q and alpha
        const long mc = req.m_control;
        RB_LAUNCH(K_NORMALS, launch_normals_f64(key, TURBO_PILOT_BASE, mc, nstreams, N, d_Z, stream));
        RB_LAUNCH(K_VOLTERRA, conv(mc, long(nstreams) * N));
        RB_LAUNCH(K_SCAN, launch_scan_f64(METHOD_RT, d_Z, long(nstreams) * N, d_Wt, mc, N, d_groups,
                                          ngroups, d_comp, d_ints, stream, 1.0));
        RB_LAUNCH(K_PAYOFF, launch_turbo_q(d_ints, mc, ngroups, d_Qg, stream));
        RB_LAUNCH(K_PAYOFF, launch_turbo_alpha(d_ints, mc, d_opts, nopt, d_Qg, d_alpha, stream));
    }
    for (long p0 = 0; p0 < req.M; p0 += chunk) {
        const long np = std::min(chunk, req.M - p0);
        if (req.gauss >= 1) {
            RB_LAUNCH(K_NORMALS, launch_normals_sobol_f64(d_sobol, d_shift, p0, np, nstreams, N, d_Z, stream));
            if (req.gauss == 2) RB_LAUNCH(K_NORMALS, launch_bridge_f64(d_Z, np * nstreams, N, stream));
        } else
            RB_LAUNCH(K_NORMALS, launch_normals_f64(key, p0, np, nstreams, N, d_Z, stream));
        RB_LAUNCH(K_VOLTERRA, conv(np, long(nstreams) * N));
        RB_LAUNCH(K_SCAN, launch_scan_f64(req.method, d_Z, long(nstreams) * N, d_Wt, np, N, d_groups,
                                          ngroups, d_comp, d_ints, stream));
        if (req.turbo) {
            RB_LAUNCH(K_SCAN, launch_scan_f64(METHOD_RT, d_Z, long(nstreams) * N, d_Wt, np, N, d_groups,
                                              ngroups, d_comp, d_ints2, stream, -1.0));
            RB_LAUNCH(K_PAYOFF, launch_turbo_payoff(d_ints, d_ints2, np, d_opts, nopt, d_Qg, d_alpha,
                                                    req.q_max, p0 / 256, nblk, d_partial, d_exc, stream));
        } else {
            RB_LAUNCH(K_PAYOFF, launch_payoff_f64(d_ints, np, p0, d_opts, nopt, req.method, p0 / 256,
                                                  nblk, d_partial, nullptr, stream));
        }
    }
    RB_LAUNCH(K_REDUCE, launch_reduce_f64(d_partial, nblk, req.M, nopt, d_price, d_stat, stream));
    if (req.turbo)
        RB_LAUNCH(K_REDUCE, launch_turbo_adjust(d_price, d_opts, nopt, d_Qg, d_alpha, req.q_max, d_exc,
                                                nblk, req.M, d_count, stream));
    RB_LAUNCH(K_IV, launch_iv_f64(d_price, d_opts, nopt, d_iv, stream));
    return ST_OK;
}

int Plan::run(std::string& err) {
    if (use_graph) {
        // first run: record the launches into a graph; every run replays it
        if (!gexec) {
            // one eager run first: kernel attributes are set outside the capture, and the
            // captured graph then replays the same launches
            RB_TRY(hipStreamSynchronize(stream));
            {
                const int rc = enqueue(err);
                if (rc) return rc;
            }
            RB_TRY(hipStreamSynchronize(stream));
            capturing = true;
            RB_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(err);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            capturing = false;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            RB_TRY(e);
            graph = g;
            RB_TRY(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        RB_TRY(hipGraphLaunch(gexec, stream));
        RB_TRY(hipStreamSynchronize(stream));
        return ST_OK;
    }
    begin_timing();
    const int rc = enqueue(err);
    if (rc) return rc;
    RB_TRY(hipStreamSynchronize(stream));
    end_timing();
    return ST_OK;
}

int Plan::run_samples(const double* Z_host, double* payoff_host, std::string& err) {
    const int N = req.N;
    const int nopt = int(lay.options.size());
    begin_timing();
    for (long p0 = 0; p0 < req.M; p0 += chunk) {
        const long np = std::min(chunk, req.M - p0);
        RB_TRY(hipMemcpyAsync(d_Z, Z_host + size_t(p0) * nstreams * N,
                              size_t(np) * nstreams * N * sizeof(double), hipMemcpyHostToDevice, stream));
        RB_LAUNCH(K_VOLTERRA, conv(np, long(nstreams) * N));
        RB_LAUNCH(K_SCAN, launch_scan_f64(req.method, d_Z, long(nstreams) * N, d_Wt, np, N, d_groups,
                                          int(lay.groups.size()), d_comp, d_ints, stream));
        RB_LAUNCH(K_PAYOFF, launch_payoff_f64(d_ints, np, 0, d_opts, nopt, req.method, p0 / 256, nblk,
                                              d_partial, d_payoff, stream));
        RB_TRY(hipMemcpyAsync(payoff_host + size_t(p0) * nopt, d_payoff, size_t(np) * nopt * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        RB_TRY(hipStreamSynchronize(stream));
    }
    RB_LAUNCH(K_REDUCE, launch_reduce_f64(d_partial, nblk, req.M, nopt, d_price, d_stat, stream));
    RB_LAUNCH(K_IV, launch_iv_f64(d_price, d_opts, nopt, d_iv, stream));
    RB_TRY(hipStreamSynchronize(stream));
    end_timing();
    return ST_OK;
}

// FFT in FP64 (iterative radix-2, sign -1 forward) for the host-side table FFT(Gamma)
void fft_host(std::vector<std::complex<double>>& a, int sign) {
    const int n = int(a.size());
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const double ang = sign * 2.0 * M_PI * k / len;
                const std::complex<double> w(std::cos(ang), std::sin(ang));
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

int Plan::init_fft(std::string& err) {
    const int N = req.N;
    log2L = 0;
    while ((1 << log2L) < 2 * N - 1) ++log2L;
    if (log2L < 1) log2L = 1;
    const int L = 1 << log2L;
    hipDeviceProp_t prop;
    RB_TRY(device_props(device, prop));
    if (volterra_fft_lds_bytes(log2L) > prop.sharedMemPerBlock) {
        err = "N too large for the LDS FFT";
        return ST_ERR_LIMIT;
    }
    // twiddles computed in FP64, stored in FP32
    std::vector<float2> tw(L / 2);
    for (int k = 0; k < L / 2; ++k) {
        const double ang = -2.0 * M_PI * double(k) / double(L);
        tw[k] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
    }
    std::vector<float2> gh(lay.units.size() * size_t(L));
    for (size_t u = 0; u < lay.units.size(); ++u) {
        std::vector<std::complex<double>> a(L, 0.0);
        for (int k = 0; k < N; ++k) a[k] = lay.gamma[u * N + k];
        fft_host(a, -1);
        for (int k = 0; k < L; ++k) gh[u * L + k] = make_float2(float(a[k].real()), float(a[k].imag()));
    }
    RB_TRY(upload(&d_tw, tw, stream));
    RB_TRY(upload(&d_ghat, gh, stream));
    return ST_OK;
}

hipError_t Plan::conv(long np, long zstride) {
    if (req.scheme == SCHEME_CHOLESKY)
        return launch_volterra_chol(d_Z, zstride, np, req.N, int(lay.units.size()), d_chol, d_Wt, stream);
    if (conv_fft)
        return launch_volterra_fft(d_Z, zstride, np, req.N, log2L, int(lay.units.size()), d_units, d_ghat,
                                   d_tw, d_Wt, stream);
    return launch_volterra_conv_f64(d_Z, zstride, np, req.N, int(lay.units.size()), d_units, d_gamma, d_Wt,
                                    lds_limit, stream);
}

int Plan::turbo_results(double* alpha, double* Q, long* exceptions, std::string& err) {
    if (!req.turbo) {
        err = "not a turbo-charged plan";
        return ST_ERR_ARG;
    }
    const int nopt = int(lay.options.size());
    if (alpha) RB_TRY(hipMemcpyAsync(alpha, d_alpha, nopt * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (exceptions)
        RB_TRY(hipMemcpyAsync(exceptions, d_count, nopt * sizeof(long), hipMemcpyDeviceToHost, stream));
    std::vector<double> qg(lay.groups.size());
    RB_TRY(hipMemcpyAsync(qg.data(), d_Qg, qg.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    RB_TRY(hipStreamSynchronize(stream));
    if (Q)
        for (int i = 0; i < nopt; ++i) Q[i] = qg[lay.options[i].group];
    return ST_OK;
}

int Plan::results(double* price, double* iv, double* stat, std::string& err) {
    if (fused || fused_fft) {
        const int rc = fused ? fused_batch->results(0, price, iv, stat, stream, err)
                             : fused_fft->results(price, iv, stat, stream, err);
        if (rc) return rc;
        RB_TRY(hipStreamSynchronize(stream));
        return ST_OK;
    }
    const size_t n = lay.options.size() * sizeof(double);
    if (price) RB_TRY(hipMemcpyAsync(price, d_price, n, hipMemcpyDeviceToHost, stream));
    if (iv) RB_TRY(hipMemcpyAsync(iv, d_iv, n, hipMemcpyDeviceToHost, stream));
    if (stat) RB_TRY(hipMemcpyAsync(stat, d_stat, n, hipMemcpyDeviceToHost, stream));
    RB_TRY(hipStreamSynchronize(stream));
    return ST_OK;
}

}  // namespace rb
