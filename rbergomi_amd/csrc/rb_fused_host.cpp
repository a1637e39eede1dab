// rb_fused_host.cpp -- host planning and launching of the fused mixed-precision kernel.
//
// plan_fused_request builds, from one request's deduplicated layout: the per-H tables of
// Toeplitz A fragments (hi + lo, or FP32), the per-H tables P_j = (j - off)^(2H), the group exponent
// coefficients, the payoff groups (group, rho), the option list and the assignment of 32-row
// output blocks to the four waves of a workgroup (balanced triangular work).
// FusedBatch concatenates any number of such requests into one launch: one table of requests,
// one list of work items (up to 1024 paths of one request each, at least 256 items per request
// where its tiles allow, the last eighth of a very large request in items an eighth of the
// size) ordered by decreasing cost and handed out through an atomic counter,
// and one ordered reduction of every request's option sums.
// A request's results depend only on its own items, so they are bit-identical alone and in any
// batch.
#include "rb_fused_host.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "rb_engine.h"
#include "rb_kernels.h"

namespace rb {

// ---------------------------------------------------------------- tile shape choice
// Path blocks of 32 per tile (the tile is 32 PB paths by Np steps; the four waves split into
// 4 / PB per path block).  Measured on MI355X (split FP16, Romano-Touzi, 205 options, M = 2^21):
// N = 96: 2.96 / 2.58 / 2.31 ms for PB = 1 / 2 / 4; N = 128: 3.04 / 2.72 / 4.39; N = 160: 3.72 /
// 3.20 / 5.39; N = 192: 3.91 / 3.33 / 6.16; N = 256: 4.27 / 6.63 / 7.29; N >= 320: PB = 1 best.
// The default follows that; autotune_fused replaces it per (Np, mode, estimator) by timing.
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 3>, int> g_tune;  // (Np, mode, mc) -> PB
std::once_flag g_tune_env;
}  // namespace

int fused_default_pb(int NB) { return NB == 1 ? 4 : NB == 2 ? 2 : NB == 3 ? 4 : NB <= 6 ? 2 : 1; }

int fused_pb_choice(int Np, int mode, int mc) {
    std::call_once(g_tune_env, [] {
        // a cache file written by tune_save, loaded once per process
        if (const char* path = std::getenv("RBERGOMI_TUNE_CACHE")) {
            std::string err;
            (void)tune_load(path, err);
        }
    });
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const auto it = g_tune.find({Np, mode, mc});
    return it != g_tune.end() ? it->second : fused_default_pb(Np / 32);
}

void tune_set(int Np, int mode, int mc, int pb) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[{Np, mode, mc}] = pb;
}

void tune_clear() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
}

// text file, one line per entry: "Np mode mc pb"
int tune_save(const char* path, std::string& err) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    FILE* f = std::fopen(path, "w");
    if (!f) {
        err = std::string("cannot write ") + path;
        return ST_ERR_ARG;
    }
    for (const auto& e : g_tune) std::fprintf(f, "%d %d %d %d\n", e.first[0], e.first[1], e.first[2], e.second);
    std::fclose(f);
    return ST_OK;
}

int tune_load(const char* path, std::string& err) {
    FILE* f = std::fopen(path, "r");
    if (!f) {
        err = std::string("cannot read ") + path;
        return ST_ERR_ARG;
    }
    std::vector<std::array<int, 4>> rows;
    int np, mode, mc, pb, line = 0;
    char buf[256];
    while (std::fgets(buf, sizeof buf, f)) {
        ++line;
        if (buf[0] == '#' || buf[0] == '\n') continue;
        if (std::sscanf(buf, "%d %d %d %d", &np, &mode, &mc, &pb) != 4 || np < 32 || np % 32 || mode < 0 ||
            mode > 2 || mc < 0 || mc > 1 || (pb != 1 && pb != 2 && pb != 4)) {
            std::fclose(f);
            err = std::string(path) + ": malformed entry on line " + std::to_string(line);
            return ST_ERR_ARG;
        }
        rows.push_back({np, mode, mc, pb});
    }
    std::fclose(f);
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto& r : rows) g_tune[{r[0], r[1], r[2]}] = r[3];
    return ST_OK;
}

std::vector<std::array<int, 4>> tune_table() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::vector<std::array<int, 4>> t;
    for (const auto& e : g_tune) t.push_back({e.first[0], e.first[1], e.first[2], e.second});
    return t;
}

static uint16_t to_bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return uint16_t(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                 // round to nearest even
    return uint16_t(u >> 16);
}
static float from_bf16_bits(uint16_t b) {
    const uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// hi + lo split of x in the operand type of the mode; returns the raw 16-bit patterns
static void split16(double x, int mode, uint16_t& hi, uint16_t& lo) {
    if (mode == 0) {
        const _Float16 h = (_Float16)(float)x;
        const _Float16 l = (_Float16)(float)(x - (double)(float)h);
        std::memcpy(&hi, &h, 2);
        std::memcpy(&lo, &l, 2);
    } else {
        hi = to_bf16_bits((float)x);
        lo = to_bf16_bits((float)(x - (double)from_bf16_bits(hi)));
    }
}

// Units, groups (sorted by unit), payoff groups (group, rho) and options (sorted by unit) of a
// request for the fused kernels, with the per-unit tables P_j = (j - off)^(2H) of Np entries,
// P_j at index j - pshift (1 <= j < N).  pgmax: payoff groups held at once; gmax: most groups of
// one unit.
void fused_tables(const Request& req, const Layout& lay, int Np, int frag_stride, int pshift, FusedHostReq& out,
                  int& pgmax, int& gmax) {
    const int nu = int(lay.units.size());
    out.fu.assign(nu, FusedUnit());
    out.fg.clear();
    out.fpg.clear();
    out.fo.clear();
    out.ptab.assign(size_t(nu) * Np, 0.f);
    const double log2e = 1.4426950408889634;
    const double off = req.comp == COMP_EXACT ? 0.0 : 1.0;
    pgmax = 1;
    gmax = 1;
    for (int u = 0; u < nu; ++u) {
        FusedUnit& U = out.fu[u];
        const double H = lay.units[u].H;
        U.b = float(lay.units[u].b);
        U.sc = float(lay.units[u].sc);
        U.frag_off = u * frag_stride;
        U.p_off = u * Np;
        for (int j = 1; j < req.N; ++j)
            out.ptab[size_t(u) * Np + j - pshift] = float(std::pow(double(j) - off, 2.0 * H));
        U.g_begin = int(out.fg.size());
        U.pg_begin = int(out.fpg.size());
        U.o_begin = int(out.fo.size());
        for (size_t g = 0; g < lay.groups.size(); ++g) {
            const Group& G = lay.groups[g];
            if (G.unit != u) continue;
            FusedGroup F;
            F.c1 = float(0.5 * G.eta * G.TH * log2e);
            // at least FLT_MIN: the kernels zero a step's term by P = +inf (al P = +inf)
            F.al = std::max(float(0.25 * G.eta * G.eta * std::pow(G.dt, 2.0 * H) * log2e), FLT_MIN);
            F.be = float(-0.5 * std::log(G.xi) * log2e);
            F.xi = G.xi;
            F.sxi = std::sqrt(G.xi);
            F.dt = G.dt;
            F.sdt = G.sdt;
            F.pg_begin = int(out.fpg.size());
            F.o_begin = int(out.fo.size());
            const int fgi = int(out.fg.size());
            for (size_t I = 0; I < lay.options.size(); ++I) {
                const Option& o = lay.options[I];
                if (o.group != int(g)) continue;
                int pg = -1;
                for (int q = F.pg_begin; q < int(out.fpg.size()); ++q)
                    if (out.fpg[q].rho == o.rho) pg = q;
                if (pg < 0) {
                    pg = int(out.fpg.size());
                    out.fpg.push_back(FusedPG{fgi, 0, o.rho});
                }
                out.fo.push_back(FusedOption{pg, int(I), o.K, o.lnK});
            }
            F.pg_end = int(out.fpg.size());
            F.o_end = int(out.fo.size());
            out.fg.push_back(F);
        }
        U.g_end = int(out.fg.size());
        U.pg_end = int(out.fpg.size());
        U.o_end = int(out.fo.size());
        // payoff groups held at once: those of one pass (at most FUSED_GREG groups), so every
        // window of FUSED_GREG groups
        for (int g0 = U.g_begin; g0 < U.g_end; ++g0) {
            const int g1 = std::min(U.g_end, g0 + FUSED_GREG) - 1;
            pgmax = std::max(pgmax, out.fg[g1].pg_end - out.fg[g0].pg_begin);
        }
        gmax = std::max(gmax, U.g_end - U.g_begin);
    }
}

int plan_fused_request(const Request& req, const Layout& lay, int mode, FusedHostReq& out,
                       std::string& err, int pb) {
    const int N = req.N;
    const int Np = (N + 31) / 32 * 32;
    if (Np > 1024) {
        err = "the split-precision kernel covers N <= 1024";
        return ST_ERR_LIMIT;
    }
    FusedReq& a = out.r;
    std::memset(&a, 0, sizeof(a));
    a.N = N;
    a.Np = Np;
    a.NK = Np / 16;
    a.chunk_u4 = fused_chunk_u4(mode);
    const int NB = Np / 32;
    // path blocks of 32 per tile: given, else the tuned or default choice for this shape
    a.PB = pb > 0 ? pb : fused_pb_choice(Np, mode, req.method == METHOD_MC ? 1 : 0);
    if (a.PB != 1 && a.PB != 2 && a.PB != 4) {
        err = "path blocks per tile must be 1, 2 or 4";
        return ST_ERR_ARG;
    }
    a.TP = 32 * a.PB;
    a.log2TP = a.PB == 1 ? 5 : (a.PB == 2 ? 6 : 7);
    // ---- waves: 4 / PB per path block; a block's output row blocks (cost rb + 1) go to its
    //      waves by longest-processing-time first
    std::vector<std::vector<int>> wrb(4);
    const int wpb = 4 / a.PB;
    for (int b = 0; b < a.PB; ++b) {
        std::vector<long> load(wpb, 0);
        for (int rb = NB - 1; rb >= 0; --rb) {
            const int w = int(std::min_element(load.begin(), load.end()) - load.begin());
            wrb[b * wpb + w].push_back(rb);
            load[w] += rb + 1;
        }
    }
    for (int w = 0; w < 4; ++w) a.wave_pb[w] = wrb[w].empty() ? -1 : w / wpb;
    size_t upw = 1;
    for (auto& v : wrb) upw = std::max(upw, v.size());
    if (upw > size_t(FUSED_MAX_UPW)) {
        err = "too many row blocks per wave for " + std::to_string(a.PB) + " path blocks per tile";
        return ST_ERR_LIMIT;
    }
    out.upw = upw <= 1 ? 1 : upw <= 2 ? 2 : upw <= 4 ? 4 : 8;
    for (int w = 0; w < 4; ++w) {
        // row blocks of a wave in decreasing order (the GEMM phases rely on it)
        std::sort(wrb[w].begin(), wrb[w].end(), [](int x, int y) { return x > y; });
        a.wave_nu[w] = int(wrb[w].size());
        for (int s = 0; s < FUSED_MAX_UPW; ++s) a.wave_rb[w][s] = s < int(wrb[w].size()) ? wrb[w][s] : -1;
    }
    int pgmax = 1, gmax = 1;
    // the GEMM kernel reads P_{j} at j - 1: a lane's four rows 8c + 4h + 0..3 are one aligned float4
    fused_tables(req, lay, Np, a.NK * 128, 1, out, pgmax, gmax);
    const int nu = int(lay.units.size());
    // rows j >= N of the last row block: P = +inf at index j - 1 and al >= FLT_MIN, so the
    // integrand exp2(c1 x - al P - be) is exactly 0 there (the kernel does not test rows)
    for (int u = 0; u < nu; ++u)
        for (int i = std::max(N - 1, 0); i < Np; ++i) out.ptab[size_t(u) * Np + i] = INFINITY;
    // the kernel reads each group's row terms al P_j + be (Np per group, the unit's groups in
    // order from p_off): one FP32 multiply-add rounded once, as the device's fmaf would give,
    // taken out of the kernel's (row, group) loop.  al P = +inf stays +inf.
    {
        std::vector<float> q;
        q.reserve(out.fg.size() * size_t(Np));
        for (int u = 0; u < nu; ++u) {
            FusedUnit& U = out.fu[u];
            const float* P = &out.ptab[size_t(u) * Np];
            U.p_off = int(q.size());
            for (int g = U.g_begin; g < U.g_end; ++g)
                for (int i = 0; i < Np; ++i) q.push_back(std::fmaf(out.fg[g].al, P[i], out.fg[g].be));
        }
        out.ptab.swap(q);
    }
    // ---- A fragments: entry (d, lane, slot j) = G[d + m - 8h - j], G[0] = a, G[t>0] = Gamma_t
    //      (0 outside [0, N)).  The table holds each G[t] about 64 times, so the operand split
    //      is computed once per t and the entries copied from it.
    out.frag.assign(size_t(nu) * a.NK * 128, uint4());
    std::vector<uint16_t> ghi(N), glo(N);
    std::vector<float> gf(N);
    for (int u = 0; u < nu; ++u) {
        const double* gam = &lay.gamma[size_t(u) * N];
        for (int t = 0; t < N; ++t) {
            const double g = t == 0 ? lay.units[u].a : gam[t];
            if (mode == 2)
                gf[t] = float(g);
            else
                split16(g, mode, ghi[t], glo[t]);
        }
        for (int f = 0; f < a.NK; ++f) {
            const int d = 16 * f - 16;
            uint4* e = &out.frag[(size_t(u) * a.NK + f) * 128];
            for (int lane = 0; lane < 64; ++lane) {
                const int m = lane & 31, hh = lane >> 5;
                uint16_t hi[8], lo[8];
                float g32[8];
                for (int j = 0; j < 8; ++j) {
                    const int t = d + m - 8 * hh - j;
                    const bool in = t >= 0 && t < N;
                    hi[j] = in ? ghi[t] : 0;
                    lo[j] = in ? glo[t] : 0;
                    g32[j] = in ? gf[t] : 0.0f;
                }
                if (mode == 2) {  // FP32 operands: slots 0..3, then 4..7
                    std::memcpy(&e[lane], g32, 16);
                    std::memcpy(&e[64 + lane], g32 + 4, 16);
                } else {
                    std::memcpy(&e[lane], hi, 16);
                    std::memcpy(&e[64 + lane], lo, 16);
                }
            }
        }
    }
    a.nunits = nu;
    a.pgmax = pgmax;
    a.out_size = int(lay.options.size());
    a.acc_in_lds = a.out_size > FUSED_MAX_REG_OPTS ? 1 : 0;
    // lanes per option in the payoff stage: the largest power of two S <= 32 with S*options <= 256
    a.log2slices = 0;
    while (!a.acc_in_lds && a.log2slices < 5 && (a.out_size << (a.log2slices + 1)) <= 256) ++a.log2slices;
    // with one lane per option (129 to 256 options) the lanes left over go to the first options,
    // two lanes each, in whole waves (32 options a wave), so that no wave prices with most of its
    // lanes idle: 205 options take 32 lane pairs and 173 single lanes
    a.pay2 = !a.acc_in_lds && a.log2slices == 0 ? std::min(a.out_size, (256 - a.out_size) / 32 * 32) : 0;
    a.mc = req.method == METHOD_MC ? 1 : 0;
    a.M = req.M;
    a.ntiles = (req.M + a.TP - 1) / a.TP;
    // work items of at most 1024 paths, and at least 256 per request where its tiles allow, so a
    // small request alone still spreads over the chip (N = 64, M = 2^16: 64 items of 1024 paths
    // took 0.170 ms, 256 items of 256 paths 0.089 ms).  The item size depends on the request
    // only, so its sums are added in the same order alone and in any batch.
    a.tiles_per_item = int(std::max<long>(1, std::min<long>(1024 / a.TP, (a.ntiles + 255) / 256)));
    a.nitems = (a.ntiles + a.tiles_per_item - 1) / a.tiles_per_item;
    a.items_big = a.nitems;
    a.tiles_small = a.tiles_per_item;
    // a request of 4096 or more items is in practice a launch of its own: its last eighth of
    // tiles goes in items of an eighth of the size, handed out last (cheapest first in the
    // batch's order), so the workgroups finish together rather than one large item apart
    // (N = 512, M = 2^22, Romano-Touzi: 7.05 -> 6.70 ms with every item of 256 paths; plain
    // Monte Carlo, with no data-dependent branch, was even without it)
    if (a.nitems >= 4096 && a.tiles_per_item >= 8) {
        a.tiles_small = a.tiles_per_item / 8;
        a.items_big = (a.ntiles - a.ntiles / 8) / a.tiles_per_item;
        const long rest = a.ntiles - a.items_big * a.tiles_per_item;
        a.nitems = a.items_big + (rest + a.tiles_small - 1) / a.tiles_small;
    }
    a.key = philox_key(req.seed.data(), int(req.seed.size()));
    a.gchunk = std::min(gmax, FUSED_GREG);
    out.gmax = gmax;
    out.ivopts = lay.options;
    // cost of one item (relative): RNG + integrals ~ N * groups, GEMM ~ N^2 / 64, payoffs ~ options
    out.item_cost = double(a.tiles_per_item * a.TP) *
                    (double(N) * (2.0 + 2.0 * a.mc + gmax) + double(Np) * Np / 64.0 * nu + 4.0 * a.out_size);
    return ST_OK;
}

#define RB_TRY(expr)                                                         \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return e_ == hipErrorOutOfMemory ? ST_ERR_NOMEM : ST_ERR_DEVICE; \
        }                                                                    \
    } while (0)

template <class T>
static hipError_t upload_vec(T** p, const std::vector<T>& v, hipStream_t st) {
    // stream-ordered allocation: freeing a batch never waits for other streams
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(p), std::max<size_t>(v.size(), 1) * sizeof(T), st);
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

template <class T>
static hipError_t alloc_async(T** p, size_t n, hipStream_t st) {
    return hipMallocAsync(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T), st);
}

// n elements of T from the arena (256-byte aligned), or a stream-ordered allocation without one
template <class T>
static hipError_t batch_alloc(DevArena* ar, T** p, size_t n, hipStream_t st) {
    if (!ar) return alloc_async(p, n, st);
    const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
    if (ar->used + bytes > ar->size) return hipErrorOutOfMemory;
    *p = reinterpret_cast<T*>(ar->base + ar->used);
    ar->used += bytes;
    return hipSuccess;
}

size_t fused_batch_lds(const std::vector<const FusedHostReq*>& reqs, size_t lds_cu, FusedBatchArgs* a, int* cap) {
    int bf = 0, wp = 0, pg = 0, ac = 0, ns = 2;
    for (const FusedHostReq* q : reqs) {
        const FusedReq& r = q->r;
        ns = r.mc ? 3 : 2;
        bf = std::max(bf, r.PB * r.NK * r.chunk_u4);
        wp = std::max(wp, 4 * r.gchunk * 32 * ns);
        pg = std::max(pg, r.TP * r.pgmax);
        ac = std::max(ac, r.acc_in_lds ? 2 * r.out_size : 0);
    }
    const size_t stat = sizeof(FusedReq) + 16;  // the kernel's static record and item index
    const size_t half = lds_cu / 2;
    const size_t rest = size_t(bf) * 16 + size_t(pg) * sizeof(PathPG) + size_t(ac) * 8 + stat;
    int c_cap = 0;
    if (rest + size_t(wp) * 8 > half) {
        for (int c = FUSED_GREG; c >= 1 && c_cap == 0; --c)
            if (rest + size_t(4 * c * 32 * ns) * 8 <= half) c_cap = c;
        if (c_cap > 0) {
            wp = 0;
            for (const FusedHostReq* q : reqs) wp = std::max(wp, 4 * std::min(q->r.gchunk, c_cap) * 32 * ns);
        }
    }
    if (a) {
        a->bf_u4 = bf;
        a->wpart_d = wp;
        a->pgd_n = pg;
        a->acc_d = ac;
    }
    if (cap) *cap = c_cap;
    return rest + size_t(wp) * 8;
}

bool fused_batch_accepts(const std::vector<const FusedHostReq*>& reqs, size_t lds_block, size_t lds_cu) {
    if (reqs.size() <= 1) return true;
    const size_t with = fused_batch_lds(reqs, lds_cu, nullptr, nullptr);
    if (with > lds_block) return false;
    if (with <= lds_cu / 2) return true;
    const std::vector<const FusedHostReq*> before(reqs.begin(), reqs.end() - 1);
    return fused_batch_lds(before, lds_cu, nullptr, nullptr) > lds_cu / 2;
}

std::vector<std::pair<int, int>> FusedBatch::split(const std::vector<const FusedHostReq*>& reqs, size_t lds_block,
                                                   size_t lds_cu, bool soft) {
    std::vector<std::pair<int, int>> runs;
    std::vector<const FusedHostReq*> cur;
    int begin = 0;
    for (int i = 0; i < int(reqs.size()); ++i) {
        cur.push_back(reqs[i]);
        const bool fits = soft ? fused_batch_accepts(cur, lds_block, lds_cu)
                               : cur.size() == 1 || fused_batch_lds(cur, lds_cu, nullptr, nullptr) <= lds_block;
        if (!fits) {
            runs.push_back({begin, i});
            begin = i;
            cur.assign(1, reqs[i]);
        }
    }
    if (!reqs.empty()) runs.push_back({begin, int(reqs.size())});
    return runs;
}

FusedBatch::~FusedBatch() { release(); }

void FusedBatch::release() {
    if (d_block && !arena) (void)hipFreeAsync(d_block, stream);
    d_block = nullptr;
    d_reqs = nullptr;
    d_fu = nullptr;
    d_fg = nullptr;
    d_fpg = nullptr;
    d_fo = nullptr;
    d_frag = nullptr;
    d_ptab = nullptr;
    d_item_req = nullptr;
    d_item_local = nullptr;
    d_counter = nullptr;
    d_partial = nullptr;
    d_price = d_stat = d_iv = nullptr;
    d_ivopts = nullptr;
    d_row_off = d_row_n = d_row_M = nullptr;
}

size_t FusedBatch::footprint(const std::vector<const FusedHostReq*>& reqs) {
    size_t b = 0;
    for (const FusedHostReq* q : reqs) {
        b += q->frag.size() * 16 + q->ptab.size() * 4 + q->fo.size() * sizeof(FusedOption) +
             q->fg.size() * sizeof(FusedGroup) + q->fpg.size() * sizeof(FusedPG) +
             q->fu.size() * sizeof(FusedUnit) + sizeof(FusedReq);
        b += size_t(q->r.out_size) * q->r.nitems * 16 + size_t(q->r.nitems) * 8;
        b += size_t(q->r.out_size) * (3 * 8 + sizeof(Option) + 24);
    }
    return b;
}

// copies v into an existing device buffer of v.size() elements
template <class T>
static hipError_t batch_refill(T* p, const std::vector<T>& v, hipStream_t st) {
    if (v.empty()) return hipSuccess;
    return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

int FusedBatch::build(const std::vector<const FusedHostReq*>& reqs, int mode_, int device,
                      hipStream_t st, std::string& err, bool refill, bool check_only) {
    if (refill && !d_reqs) {
        err = "no batch to refill";
        return ST_ERR_ARG;
    }
    if (!refill) release();
    // a refill that fails its check, or only checks, leaves the batch as it was
    struct Saved {
        FusedBatch& b;
        int mode, nreq, upw, greg, total_out;
        bool mc;
        hipStream_t stream;
        std::vector<long> out_off;
        FusedBatchArgs args;
        bool keep = false;
        ~Saved() {
            if (keep) return;
            b.mode = mode;
            b.nreq = nreq;
            b.upw = upw;
            b.greg = greg;
            b.total_out = total_out;
            b.mc = mc;
            b.stream = stream;
            b.out_off = out_off;
            b.args = args;
        }
    } saved{*this, mode, nreq, upw, greg, total_out, mc, stream, out_off, args};
    saved.keep = !refill;
    stream = st;
    mode = mode_;
    nreq = int(reqs.size());
    upw = 1;
    // concatenate the tables, fixing the offsets inside each request
    std::vector<FusedReq> R;
    std::vector<FusedUnit> fu;
    std::vector<FusedGroup> fg;
    std::vector<FusedPG> fpg;
    std::vector<FusedOption> fo;
    std::vector<uint4> frag;
    std::vector<float> ptab;
    std::vector<Option> ivo;
    std::vector<long> row_off, row_n, row_M;
    out_off.assign(nreq + 1, 0);
    long poff = 0;
    for (int q = 0; q < nreq; ++q) {
        const FusedHostReq& H = *reqs[q];
        FusedReq r = H.r;
        upw = std::max(upw, H.upw);
        const int u0 = int(fu.size()), g0 = int(fg.size()), pg0 = int(fpg.size()), o0 = int(fo.size());
        const int f0 = int(frag.size()), p0 = int(ptab.size());
        for (FusedUnit U : H.fu) {
            U.frag_off += f0;
            U.p_off += p0;
            U.g_begin += 0;
            fu.push_back(U);
        }
        for (const FusedGroup& G : H.fg) fg.push_back(G);
        for (FusedPG P : H.fpg) {
            P.group += 0;  // group indices are request-local
            fpg.push_back(P);
        }
        for (const FusedOption& O : H.fo) fo.push_back(O);
        frag.insert(frag.end(), H.frag.begin(), H.frag.end());
        ptab.insert(ptab.end(), H.ptab.begin(), H.ptab.end());
        // pointers are fixed after upload; keep the element offsets for now
        r.units = reinterpret_cast<const FusedUnit*>(size_t(u0));
        r.groups = reinterpret_cast<const FusedGroup*>(size_t(g0));
        r.pgs = reinterpret_cast<const FusedPG*>(size_t(pg0));
        r.opts = reinterpret_cast<const FusedOption*>(size_t(o0));
        r.partial = reinterpret_cast<double*>(size_t(poff));
        for (int I = 0; I < r.out_size; ++I) {
            row_off.push_back(poff + long(I) * r.nitems * 2);
            row_n.push_back(r.nitems);
            row_M.push_back(r.M);
        }
        poff += long(r.out_size) * r.nitems * 2;
        out_off[q + 1] = out_off[q] + r.out_size;
        ivo.insert(ivo.end(), H.ivopts.begin(), H.ivopts.end());
        if (q > 0 && r.mc != R[0].mc) {
            err = "a batch holds requests of one estimator (Romano-Touzi or plain Monte Carlo)";
            return ST_ERR_ARG;
        }
        R.push_back(r);
    }
    mc = nreq > 0 && R[0].mc != 0;
    {
        int gm = 1;
        for (int q = 0; q < nreq; ++q) gm = std::max(gm, reqs[q]->gmax);
        greg = fused_greg(gm);
    }
    total_out = int(out_off[nreq]);
    hipDeviceProp_t prop;
    RB_TRY(device_props(device, prop));
    if (!arena) keep_pool_cached(device);
    // Two workgroups per CU need at most half the CU's LDS each, the static request record
    // included; fused_batch_lds cuts the cross-wave partials' chunk of groups when that helps.
    const size_t lds_cu = size_t(prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                        : prop.sharedMemPerBlock);
    int cap = 0;
    const size_t lds_total = fused_batch_lds(reqs, lds_cu, &args, &cap);
    if (cap > 0)
        for (FusedReq& r : R) r.gchunk = std::min(r.gchunk, cap);
    if (lds_total > prop.sharedMemPerBlock) {
        err = "fused kernel needs " + std::to_string(lds_total) + " bytes of LDS (too many options or groups; "
              "a request that fits alone may not fit with others in one launch)";
        return ST_ERR_LIMIT;
    }
    // work items, most expensive first
    std::vector<std::pair<double, std::pair<int, int>>> items;
    for (int q = 0; q < nreq; ++q)
        for (long i = 0; i < R[q].nitems; ++i)
            items.push_back({i < R[q].items_big ? reqs[q]->item_cost
                                                : reqs[q]->item_cost * R[q].tiles_small / R[q].tiles_per_item,
                             {q, int(i)}});
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<double, std::pair<int, int>>& x,
                        const std::pair<double, std::pair<int, int>>& y) { return x.first > y.first; });
    std::vector<int> item_req(items.size()), item_local(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        item_req[i] = items[i].second.first;
        item_local[i] = items[i].second.second;
    }
    args.total_items = long(items.size());
    const std::vector<size_t> shape = {fu.size(), fg.size(), fpg.size(), fo.size(), frag.size(), ptab.size(),
                                       ivo.size(), row_off.size(), item_req.size(), size_t(poff),
                                       size_t(total_out), size_t(mode_), size_t(upw), size_t(greg), size_t(mc),
                                       size_t(args.bf_u4), size_t(args.wpart_d), size_t(args.pgd_n),
                                       size_t(args.acc_d)};
    if (refill) {
        // new parameters of the same shapes into the same buffers: the request records keep their
        // pointers, so launches recorded in a graph stay valid
        if (shape != shape_) {
            err = "the new parameters change the batch's table layout (distinct H, maturities, rho, options "
                  "or tile shapes); build a new batch";
            return ST_ERR_ARG;
        }
        if (check_only) return ST_OK;
        saved.keep = true;
        RB_TRY(batch_refill(d_fu, fu, st));
        RB_TRY(batch_refill(d_fg, fg, st));
        RB_TRY(batch_refill(d_fpg, fpg, st));
        RB_TRY(batch_refill(d_fo, fo, st));
        RB_TRY(batch_refill(d_frag, frag, st));
        RB_TRY(batch_refill(d_ptab, ptab, st));
        RB_TRY(batch_refill(d_ivopts, ivo, st));
        RB_TRY(batch_refill(d_row_off, row_off, st));
        RB_TRY(batch_refill(d_row_n, row_n, st));
        RB_TRY(batch_refill(d_row_M, row_M, st));
        RB_TRY(batch_refill(d_item_req, item_req, st));
        RB_TRY(batch_refill(d_item_local, item_local, st));
        for (FusedReq& r : R) {
            r.units = d_fu + size_t(r.units);
            r.groups = d_fg + size_t(r.groups);
            r.pgs = d_fpg + size_t(r.pgs);
            r.opts = d_fo + size_t(r.opts);
            r.frag = d_frag;
            r.ptab = d_ptab;
            r.partial = d_partial + size_t(r.partial);
        }
        RB_TRY(batch_refill(d_reqs, R, st));
        RB_TRY(hipStreamSynchronize(st));
        return ST_OK;
    }
    shape_ = shape;
    // One device block: the tables, request records and item counter (one upload), then the
    // partial sums and the results price | iv | stat, adjacent so that one copy brings them back.
    // Pieces start on 256-byte boundaries.
    size_t off = 0;
    auto place = [&](size_t bytes) {
        const size_t o = off;
        off += (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        return o;
    };
    const size_t o_fu = place(fu.size() * sizeof(FusedUnit)), o_fg = place(fg.size() * sizeof(FusedGroup)),
                 o_fpg = place(fpg.size() * sizeof(FusedPG)), o_fo = place(fo.size() * sizeof(FusedOption)),
                 o_frag = place(frag.size() * sizeof(uint4)), o_ptab = place(ptab.size() * sizeof(float)),
                 o_ivo = place(ivo.size() * sizeof(Option)), o_roff = place(row_off.size() * sizeof(long)),
                 o_rn = place(row_n.size() * sizeof(long)), o_rm = place(row_M.size() * sizeof(long)),
                 o_ireq = place(item_req.size() * sizeof(int)), o_iloc = place(item_local.size() * sizeof(int)),
                 o_reqs = place(R.size() * sizeof(FusedReq)), o_cnt = place(sizeof(int));
    const size_t up_bytes = off;
    const size_t o_part = place(size_t(poff) * sizeof(double)), o_out = place(size_t(total_out) * 3 * sizeof(double));
    RB_TRY(batch_alloc(arena, &d_block, off, st));
    d_fu = reinterpret_cast<FusedUnit*>(d_block + o_fu);
    d_fg = reinterpret_cast<FusedGroup*>(d_block + o_fg);
    d_fpg = reinterpret_cast<FusedPG*>(d_block + o_fpg);
    d_fo = reinterpret_cast<FusedOption*>(d_block + o_fo);
    d_frag = reinterpret_cast<uint4*>(d_block + o_frag);
    d_ptab = reinterpret_cast<float*>(d_block + o_ptab);
    d_ivopts = reinterpret_cast<Option*>(d_block + o_ivo);
    d_row_off = reinterpret_cast<long*>(d_block + o_roff);
    d_row_n = reinterpret_cast<long*>(d_block + o_rn);
    d_row_M = reinterpret_cast<long*>(d_block + o_rm);
    d_item_req = reinterpret_cast<int*>(d_block + o_ireq);
    d_item_local = reinterpret_cast<int*>(d_block + o_iloc);
    d_reqs = reinterpret_cast<FusedReq*>(d_block + o_reqs);
    d_counter = reinterpret_cast<int*>(d_block + o_cnt);
    counter_dirty = false;
    d_partial = reinterpret_cast<double*>(d_block + o_part);
    d_price = reinterpret_cast<double*>(d_block + o_out);
    d_iv = d_price + total_out;
    d_stat = d_price + 2 * size_t(total_out);
    for (FusedReq& r : R) {
        r.units = d_fu + size_t(r.units);
        r.groups = d_fg + size_t(r.groups);
        r.pgs = d_fpg + size_t(r.pgs);
        r.opts = d_fo + size_t(r.opts);
        r.frag = d_frag;
        r.ptab = d_ptab;
        r.partial = d_partial + size_t(r.partial);
    }
    unsigned char* h = stage ? stage(up_bytes) : nullptr;
    const bool staged = h != nullptr;
    if (staged) {
        // packed into the caller's pinned staging: one copy, not waited for
        auto pack = [&](size_t o, const auto& v) {
            if (!v.empty()) std::memcpy(h + o, v.data(), v.size() * sizeof(v[0]));
        };
        pack(o_fu, fu);
        pack(o_fg, fg);
        pack(o_fpg, fpg);
        pack(o_fo, fo);
        pack(o_frag, frag);
        pack(o_ptab, ptab);
        pack(o_ivo, ivo);
        pack(o_roff, row_off);
        pack(o_rn, row_n);
        pack(o_rm, row_M);
        pack(o_ireq, item_req);
        pack(o_iloc, item_local);
        pack(o_reqs, R);
        std::memset(h + o_cnt, 0, sizeof(int));
        RB_TRY(hipMemcpyAsync(d_block, h, up_bytes, hipMemcpyHostToDevice, st));
    } else {
        // straight from the host tables (pageable memory: a packed copy would cost a host copy of
        // every table); the first launch zeroes the counter
        auto put = [&](size_t o, const auto& v) {
            return v.empty() ? hipSuccess
                             : hipMemcpyAsync(d_block + o, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, st);
        };
        RB_TRY(put(o_fu, fu));
        RB_TRY(put(o_fg, fg));
        RB_TRY(put(o_fpg, fpg));
        RB_TRY(put(o_fo, fo));
        RB_TRY(put(o_frag, frag));
        RB_TRY(put(o_ptab, ptab));
        RB_TRY(put(o_ivo, ivo));
        RB_TRY(put(o_roff, row_off));
        RB_TRY(put(o_rn, row_n));
        RB_TRY(put(o_rm, row_M));
        RB_TRY(put(o_ireq, item_req));
        RB_TRY(put(o_iloc, item_local));
        RB_TRY(put(o_reqs, R));
        counter_dirty = true;
    }
    args.reqs = d_reqs;
    args.item_req = d_item_req;
    args.item_local = d_item_local;
    args.counter = d_counter;
    const int per_cu = int(std::max<size_t>(1, std::min<size_t>(2, lds_cu / lds_total)));
    grid = int(std::min<long>(args.total_items, long(prop.multiProcessorCount) * per_cu));
    // pageable sources: the host tables die with this call
    if (!staged) RB_TRY(hipStreamSynchronize(st));
    return ST_OK;
}

int FusedBatch::launch(hipStream_t st, std::string& err, Plan* timing, hipEvent_t kbeg, hipEvent_t kend) {
    auto timed = [&](int id) { return timing ? timing->timed(id) : -1; };
    auto mark = [&](int slot) {
        if (timing) timing->mark(slot, nullptr, nullptr);
    };
    int s;
    // the upload and every complete launch leave the item counter at zero (the reduction resets
    // it); a launch that failed part-way may not have
    if (counter_dirty) RB_TRY(launch_zero_int(d_counter, 1, st));
    counter_dirty = true;
    if (kbeg) RB_TRY(hipEventRecord(kbeg, st));
    s = timed(K_FUSED);
    RB_TRY(launch_fused(mode, upw, mc, greg, args, grid, st));
    mark(s);
    if (kend) RB_TRY(hipEventRecord(kend, st));
    s = timed(K_REDUCE);
    RB_TRY(launch_reduce_rows_iv(d_partial, d_row_off, d_row_n, d_row_M, d_ivopts, total_out, d_price, d_stat, d_iv,
                                 d_counter, st));
    mark(s);
    counter_dirty = total_out < 1;
    return ST_OK;
}

int FusedBatch::results(int q, double* price, double* iv, double* stat, hipStream_t st,
                        std::string& err) {
    const size_t n = size_t(out_off[q + 1] - out_off[q]) * sizeof(double);
    if (price) RB_TRY(hipMemcpyAsync(price, d_price + out_off[q], n, hipMemcpyDeviceToHost, st));
    if (iv) RB_TRY(hipMemcpyAsync(iv, d_iv + out_off[q], n, hipMemcpyDeviceToHost, st));
    if (stat) RB_TRY(hipMemcpyAsync(stat, d_stat + out_off[q], n, hipMemcpyDeviceToHost, st));
    return ST_OK;
}

// ---------------------------------------------------------------- long paths: FFT kernel
FftFused::~FftFused() { release(); }

void FftFused::release() {
    void* bufs[] = {d_fu, d_fg, d_fpg, d_fo, d_ptab, d_ghat, d_tw, d_partial, d_counter, d_row_off, d_row_n,
                    d_row_M, d_price, d_stat, d_iv, d_ivopts};
    for (void* b : bufs)
        if (b) (void)hipFreeAsync(b, stream);
    d_fu = nullptr;
    d_fg = nullptr;
    d_fpg = nullptr;
    d_fo = nullptr;
    d_ptab = nullptr;
    d_ghat = d_tw = nullptr;
    d_partial = nullptr;
    d_counter = nullptr;
    d_row_off = d_row_n = d_row_M = nullptr;
    d_price = d_stat = d_iv = nullptr;
    d_ivopts = nullptr;
}

// P table of the long-path kernel: P_i at index i for the 1024 QPT steps its threads visit
// (QPT quads of four steps per thread, QPT of the instantiation that runs N), +inf at i = 0 and
// at i >= N so those terms vanish
static void fft_ptab(const Request& req, const Layout& lay, FusedHostReq& tabs, int& pgmax, int& gmax) {
    const int N = req.N, len = 1024 * fused_fft_qpt(N);
    fused_tables(req, lay, len, 0, 0, tabs, pgmax, gmax);
    for (size_t u = 0; u < lay.units.size(); ++u) {
        tabs.ptab[u * len] = INFINITY;
        for (int i = N; i < len; ++i) tabs.ptab[u * len + i] = INFINITY;
    }
}

int FftFused::plan(const Request& req, const Layout& lay, size_t lds_limit, size_t& footprint, std::string& err) {
    const int N = req.N;
    if (N <= 1024 || ((N + 3) / 4 + 255) / 256 > 8) {
        err = "the long-path kernel covers 1024 < N <= 8192";
        return ST_ERR_LIMIT;
    }
    int log2L = 10;  // L >= 2N - 1, at least 2^10 (the kernel's shortest transform)
    while ((1 << log2L) < 2 * N - 1) ++log2L;
    const size_t L = size_t(1) << log2L;
    FusedHostReq tabs;
    int pgmax = 1, gmax = 1;
    fft_ptab(req, lay, tabs, pgmax, gmax);
    FftFusedArgs a{};
    a.log2L = log2L;
    a.out_size = int(lay.options.size());
    a.pgu = 1;
    for (const FusedUnit& U : tabs.fu) a.pgu = std::max(a.pgu, U.pg_end - U.pg_begin);
    if (fused_fft_lds_bytes(a) > lds_limit) {
        err = "long-path kernel needs " + std::to_string(fused_fft_lds_bytes(a)) +
              " bytes of LDS (N or the option count too large)";
        return ST_ERR_LIMIT;
    }
    const long nitems = ((req.M + 1) / 2 + 31) / 32;
    footprint = lay.units.size() * L * sizeof(float2) + L / 2 * sizeof(float2) + tabs.ptab.size() * sizeof(float) +
                tabs.fo.size() * sizeof(FusedOption) + tabs.fpg.size() * sizeof(FusedPG) +
                tabs.fg.size() * sizeof(FusedGroup) + tabs.fu.size() * sizeof(FusedUnit) +
                lay.options.size() * (sizeof(Option) + 3 * sizeof(double) + 3 * sizeof(long) + size_t(nitems) * 16) + 16;
    return ST_OK;
}

int FftFused::build(const Request& req, const Layout& lay, int device, hipStream_t st, std::string& err) {
    release();
    stream = st;
    const int N = req.N;
    int log2L = 10;  // L >= 2N - 1, at least 2^10 (the kernel's shortest transform)
    while ((1 << log2L) < 2 * N - 1) ++log2L;
    const int L = 1 << log2L;
    if (((N + 3) / 4 + 255) / 256 > 8) {
        err = "the long-path kernel covers N <= 8192";
        return ST_ERR_LIMIT;
    }
    FusedHostReq tabs;
    int pgmax = 1, gmax = 1;
    fft_ptab(req, lay, tabs, pgmax, gmax);
    int pgu = 1;
    for (const FusedUnit& U : tabs.fu) pgu = std::max(pgu, U.pg_end - U.pg_begin);
    // twiddles and FFT(a, Gamma_1, ..., Gamma_{N-1}) per unit, in FP64, stored in FP32
    std::vector<float2> tw(L / 2);
    for (int k = 0; k < L / 2; ++k) {
        const double ang = -2.0 * M_PI * double(k) / double(L);
        tw[k] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
    }
    std::vector<float2> gh(lay.units.size() * size_t(L));
    for (size_t u = 0; u < lay.units.size(); ++u) {
        std::vector<std::complex<double>> g(L, 0.0);
        g[0] = lay.units[u].a;
        for (int k = 1; k < N; ++k) g[k] = lay.gamma[u * N + k];
        fft_host(g, -1);
        // bit-reversed order: the spectrum's order between the device transforms
        for (int k = 0; k < L; ++k) {
            const int b = bit_reverse(k, log2L);
            gh[u * L + k] = make_float2(float(g[b].real()), float(g[b].imag()));
        }
    }
    FftFusedArgs& a = args;
    a = FftFusedArgs{};
    a.N = N;
    a.log2L = log2L;
    a.nunits = int(lay.units.size());
    a.out_size = int(lay.options.size());
    a.mc = req.method == METHOD_MC ? 1 : 0;
    a.pgu = pgu;
    a.pairs_per_item = 32;
    a.M = req.M;
    a.npairs = (req.M + 1) / 2;
    a.nitems = (a.npairs + a.pairs_per_item - 1) / a.pairs_per_item;
    a.key = philox_key(req.seed.data(), int(req.seed.size()));
    total_out = a.out_size;
    hipDeviceProp_t prop;
    RB_TRY(device_props(device, prop));
    keep_pool_cached(device);
    const size_t lds = fused_fft_lds_bytes(a);
    if (lds > prop.sharedMemPerBlock) {
        err = "long-path kernel needs " + std::to_string(lds) + " bytes of LDS (N or the option count too large)";
        return ST_ERR_LIMIT;
    }
    std::vector<long> row_off(a.out_size), row_n(a.out_size, a.nitems), row_M(a.out_size, a.M);
    for (int I = 0; I < a.out_size; ++I) row_off[I] = long(I) * a.nitems * 2;
    RB_TRY(upload_vec(&d_fu, tabs.fu, st));
    RB_TRY(upload_vec(&d_fg, tabs.fg, st));
    RB_TRY(upload_vec(&d_fpg, tabs.fpg, st));
    RB_TRY(upload_vec(&d_fo, tabs.fo, st));
    RB_TRY(upload_vec(&d_ptab, tabs.ptab, st));
    RB_TRY(upload_vec(&d_ghat, gh, st));
    RB_TRY(upload_vec(&d_tw, tw, st));
    RB_TRY(upload_vec(&d_row_off, row_off, st));
    RB_TRY(upload_vec(&d_row_n, row_n, st));
    RB_TRY(upload_vec(&d_row_M, row_M, st));
    RB_TRY(upload_vec(&d_ivopts, lay.options, st));
    RB_TRY(alloc_async(&d_partial, size_t(a.out_size) * a.nitems * 2, st));
    RB_TRY(alloc_async(&d_counter, 1, st));
    RB_TRY(alloc_async(&d_price, size_t(a.out_size), st));
    RB_TRY(alloc_async(&d_stat, size_t(a.out_size), st));
    RB_TRY(alloc_async(&d_iv, size_t(a.out_size), st));
    a.units = d_fu;
    a.groups = d_fg;
    a.pgs = d_fpg;
    a.opts = d_fo;
    a.ptab = d_ptab;
    a.ghat = d_ghat;
    a.tw = d_tw;
    a.partial = d_partial;
    a.counter = d_counter;
    const size_t lds_cu = size_t(prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                        : prop.sharedMemPerBlock);
    const int per_cu = int(std::max<size_t>(1, std::min<size_t>(4, lds_cu / lds)));
    grid = int(std::min<long>(a.nitems, long(prop.multiProcessorCount) * per_cu));
    RB_TRY(hipStreamSynchronize(st));
    return ST_OK;
}

int FftFused::launch(hipStream_t st, std::string& err, Plan* timing, hipEvent_t kbeg, hipEvent_t kend) {
    auto timed = [&](int id) { return timing ? timing->timed(id) : -1; };
    auto mark = [&](int slot) {
        if (timing) timing->mark(slot, nullptr, nullptr);
    };
    int s;
    RB_TRY(launch_zero_int(d_counter, 1, st));
    if (kbeg) RB_TRY(hipEventRecord(kbeg, st));
    s = timed(K_FUSED);
    RB_TRY(launch_fused_fft(args, grid, st));
    mark(s);
    if (kend) RB_TRY(hipEventRecord(kend, st));
    s = timed(K_REDUCE);
    RB_TRY(launch_reduce_rows(d_partial, d_row_off, d_row_n, d_row_M, total_out, d_price, d_stat, st));
    mark(s);
    s = timed(K_IV);
    RB_TRY(launch_iv_f64(d_price, d_ivopts, total_out, d_iv, st));
    mark(s);
    return ST_OK;
}

int FftFused::results(double* price, double* iv, double* stat, hipStream_t st, std::string& err) {
    const size_t n = size_t(total_out) * sizeof(double);
    if (price) RB_TRY(hipMemcpyAsync(price, d_price, n, hipMemcpyDeviceToHost, st));
    if (iv) RB_TRY(hipMemcpyAsync(iv, d_iv, n, hipMemcpyDeviceToHost, st));
    if (stat) RB_TRY(hipMemcpyAsync(stat, d_stat, n, hipMemcpyDeviceToHost, st));
    return ST_OK;
}

// ---------------------------------------------------------------- autotuning
int autotune_fused(const Request& shape, const std::vector<int>& Ns, int mode, int device,
                   std::vector<int>& chosen, std::vector<double>& best_ms, std::string& err) {
    chosen.assign(Ns.size(), 0);
    best_ms.assign(Ns.size(), 0.0);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ST_OK;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        err = std::string(what) + ": " + hipGetErrorString(e);
        rc = e == hipErrorOutOfMemory ? ST_ERR_NOMEM : ST_ERR_DEVICE;
        return false;
    };
    if (!hip_ok(hipSetDevice(device), "hipSetDevice") ||
        !hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") ||
        !hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate"))
        goto done;
    for (size_t n = 0; n < Ns.size() && rc == ST_OK; ++n) {
        Request r = shape;
        r.N = Ns[n];
        rc = validate(r, err);
        if (rc) break;
        const Layout lay = make_layout(r);
        const int Np = (r.N + 31) / 32 * 32;
        double best = 0.0;
        int best_pb = 0;
        for (int pb : {1, 2, 4}) {
            FusedHostReq fh;
            std::string e2;
            if (plan_fused_request(r, lay, mode, fh, e2, pb) != ST_OK) continue;  // shape not valid
            FusedBatch B;
            if (B.build({&fh}, mode, device, st, e2) != ST_OK) continue;          // e.g. LDS
            // one untimed launch, then the fastest of three
            double ms = 0.0;
            for (int it = 0; it < 4 && rc == ST_OK; ++it) {
                if (!hip_ok(hipEventRecord(e0, st), "hipEventRecord")) break;
                if ((rc = B.launch(st, err)) != ST_OK) break;
                if (!hip_ok(hipEventRecord(e1, st), "hipEventRecord") ||
                    !hip_ok(hipEventSynchronize(e1), "hipEventSynchronize"))
                    break;
                float t = 0.f;
                if (!hip_ok(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime")) break;
                if (it > 0 && (ms == 0.0 || t < ms)) ms = t;
            }
            B.release();
            (void)hipStreamSynchronize(st);
            if (rc != ST_OK) break;
            if (best_pb == 0 || ms < best) {
                best = ms;
                best_pb = pb;
            }
        }
        if (rc != ST_OK) break;
        if (best_pb == 0) {
            err = "no tile shape of the fused kernel fits N = " + std::to_string(r.N);
            rc = ST_ERR_LIMIT;
            break;
        }
        tune_set(Np, mode, r.method == METHOD_MC ? 1 : 0, best_pb);
        chosen[n] = best_pb;
        best_ms[n] = best;
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

// ---------------------------------------------------------------- single-request plan
int Plan::init_fused(std::string& err) {
    fused_req.reset(new FusedHostReq());
    int rc = plan_fused_request(req, lay, fused_mode, *fused_req, err);
    if (rc) return rc;
    fused_batch.reset(new FusedBatch());
    rc = fused_batch->build({fused_req.get()}, fused_mode, device, stream, err);
    if (rc) return rc;
    nblk = fused_req->r.nitems;
    chunk = req.M;
    return ST_OK;
}

}  // namespace rb
