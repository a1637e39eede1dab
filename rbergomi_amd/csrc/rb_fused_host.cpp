// rb_fused_host.cpp -- host planning of the fused mixed-precision kernel.
//
// Builds, from the deduplicated layout: the per-H tables of Toeplitz A fragments (hi + lo),
// the per-group exp2 coefficient tables, the option lists per H and the assignment of
// 32-row output blocks to the four waves of a workgroup (balanced triangular work).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rb_engine.h"
#include "rb_fused.h"
#include "rb_kernels.h"

namespace rb {

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
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(v.size(), 1) * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

int Plan::init_fused(std::string& err) {
    const int N = req.N;
    if (req.method != METHOD_RT) {
        err = "the split-precision kernel prices the Romano-Touzi estimator only";
        return ST_ERR_ARG;
    }
    const int Np = (N + 31) / 32 * 32;
    if (Np > 1024) {
        err = "the split-precision kernel covers N <= 1024";
        return ST_ERR_LIMIT;
    }
    FusedArgs& a = fa;
    std::memset(&a, 0, sizeof(a));
    a.N = N;
    a.Np = Np;
    a.NK = Np / 16;
    const int NB = Np / 32;
    a.PB = NB >= 3 ? 1 : (NB == 2 ? 2 : 4);
    a.TP = 32 * a.PB;
    a.log2TP = a.PB == 1 ? 5 : (a.PB == 2 ? 6 : 7);
    // ---- waves: output row blocks (cost rb+1), all row blocks of a wave in one path block
    std::vector<std::vector<int>> wrb(4);
    for (int w = 0; w < 4; ++w) a.wave_pb[w] = -1;
    if (a.PB == 1) {
        std::vector<long> load(4, 0);
        for (int rb = NB - 1; rb >= 0; --rb) {
            const int w = int(std::min_element(load.begin(), load.end()) - load.begin());
            wrb[w].push_back(rb);
            load[w] += rb + 1;
        }
        for (int w = 0; w < 4; ++w)
            if (!wrb[w].empty()) a.wave_pb[w] = 0;
    } else {
        for (int w = 0; w < 4; ++w) {
            a.wave_pb[w] = w / NB;
            wrb[w].push_back(w % NB);
        }
    }
    size_t upw = 1;
    for (auto& v : wrb) upw = std::max(upw, v.size());
    fused_upw = upw <= 1 ? 1 : upw <= 2 ? 2 : upw <= 4 ? 4 : 8;
    for (int w = 0; w < 4; ++w) {
        // row blocks of a wave in decreasing order (the GEMM phases rely on it)
        std::sort(wrb[w].begin(), wrb[w].end(), [](int x, int y) { return x > y; });
        a.wave_nu[w] = int(wrb[w].size());
        for (int s = 0; s < FUSED_MAX_UPW; ++s) a.wave_rb[w][s] = s < int(wrb[w].size()) ? wrb[w][s] : -1;
    }
    // ---- units, groups (sorted by unit), payoff groups (group, rho), options (sorted by unit)
    const int nu = int(lay.units.size());
    std::vector<FusedUnit> fu(nu);
    std::vector<FusedGroup> fg;
    std::vector<FusedPG> fpg;
    std::vector<FusedOption> fo;
    std::vector<float> ptab(size_t(nu) * Np, 0.f);
    const double log2e = 1.4426950408889634;
    const double off = req.comp == COMP_EXACT ? 0.0 : 1.0;
    int gmax = 1, pgmax = 1;
    for (int u = 0; u < nu; ++u) {
        FusedUnit& U = fu[u];
        const double H = lay.units[u].H;
        U.b = float(lay.units[u].b);
        U.sc = float(lay.units[u].sc);
        U.frag_off = u * a.NK * 128;
        U.p_off = u * Np;
        for (int j = 1; j < N; ++j) ptab[size_t(u) * Np + j] = float(std::pow(double(j) - off, 2.0 * H));
        U.g_begin = int(fg.size());
        U.pg_begin = int(fpg.size());
        U.o_begin = int(fo.size());
        for (size_t g = 0; g < lay.groups.size(); ++g) {
            const Group& G = lay.groups[g];
            if (G.unit != u) continue;
            FusedGroup F;
            F.c1 = float(0.5 * G.eta * G.TH * log2e);
            F.al = float(0.25 * G.eta * G.eta * std::pow(G.dt, 2.0 * H) * log2e);
            F.be = float(-0.5 * std::log(G.xi) * log2e);
            F.pad = 0;
            F.xi = G.xi;
            F.sxi = std::sqrt(G.xi);
            F.dt = G.dt;
            F.sdt = G.sdt;
            F.pg_begin = int(fpg.size());
            const int fgi = int(fg.size());
            // payoff groups: distinct rho among this group's options, then its options
            for (size_t I = 0; I < lay.options.size(); ++I) {
                const Option& o = lay.options[I];
                if (o.group != int(g)) continue;
                int pg = -1;
                for (int q = F.pg_begin; q < int(fpg.size()); ++q)
                    if (fpg[q].rho == o.rho) pg = q;
                if (pg < 0) {
                    pg = int(fpg.size());
                    fpg.push_back(FusedPG{fgi, 0, o.rho});
                }
                fo.push_back(FusedOption{pg, int(I), o.K, o.lnK});
            }
            F.pg_end = int(fpg.size());
            fg.push_back(F);
        }
        U.g_end = int(fg.size());
        U.pg_end = int(fpg.size());
        U.o_end = int(fo.size());
        gmax = std::max(gmax, U.g_end - U.g_begin);
        pgmax = std::max(pgmax, U.pg_end - U.pg_begin);
    }
    // ---- A fragments: entry (d, lane, slot j) = G[d + m - 8h - j], G[0] = a, G[t>0] = Gamma_t
    std::vector<uint4> frag(size_t(nu) * a.NK * 128);
    for (int u = 0; u < nu; ++u) {
        const double* gam = &lay.gamma[size_t(u) * N];
        const double diag = lay.units[u].a;
        for (int f = 0; f < a.NK; ++f) {
            const int d = 16 * f - 16;
            for (int lane = 0; lane < 64; ++lane) {
                const int m = lane & 31, hh = lane >> 5;
                uint16_t hi[8], lo[8];
                for (int j = 0; j < 8; ++j) {
                    const int t = d + m - 8 * hh - j;
                    const double g = t < 0 ? 0.0 : (t == 0 ? diag : (t < N ? gam[t] : 0.0));
                    split16(g, fused_mode, hi[j], lo[j]);
                }
                uint4* e = &frag[(size_t(u) * a.NK + f) * 128];
                std::memcpy(&e[lane], hi, 16);
                std::memcpy(&e[64 + lane], lo, 16);
            }
        }
    }
    a.nunits = nu;
    a.gmax = gmax;
    a.pgmax = pgmax;
    a.out_size = int(lay.options.size());
    a.acc_in_lds = a.out_size > FUSED_MAX_REG_OPTS ? 1 : 0;
    a.M = req.M;
    a.ntiles = (req.M + a.TP - 1) / a.TP;
    a.tiles_per_item = std::max(1, 1024 / a.TP);
    a.nitems = (a.ntiles + a.tiles_per_item - 1) / a.tiles_per_item;
    a.key = philox_key(req.seed.data(), int(req.seed.size()));
    hipDeviceProp_t prop;
    RB_TRY(hipGetDeviceProperties(&prop, device));
    const size_t lds_cu = size_t(prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                        : prop.sharedMemPerBlock);
    // cross-wave group chunk: as many groups as fit two workgroups per CU, else one workgroup
    a.gchunk = 1;
    const size_t two = lds_cu / 2;
    for (int c = std::min(gmax, FUSED_GREG); c >= 1; --c) {
        a.gchunk = c;
        if (fused_lds_bytes(a) <= two) break;
    }
    if (fused_lds_bytes(a) > two)
        for (int c = std::min(gmax, FUSED_GREG); c >= 1; --c) {
            a.gchunk = c;
            if (fused_lds_bytes(a) <= prop.sharedMemPerBlock) break;
        }
    const size_t lds = fused_lds_bytes(a);
    if (lds > prop.sharedMemPerBlock) {
        err = "fused kernel needs " + std::to_string(lds) + " bytes of LDS (too many options or groups)";
        return ST_ERR_LIMIT;
    }
    const int per_cu = int(std::max<size_t>(1, std::min<size_t>(2, lds_cu / lds)));
    fused_grid = int(std::min<long>(a.nitems, long(prop.multiProcessorCount) * per_cu));
    RB_TRY(upload_vec(&d_fu, fu, stream));
    RB_TRY(upload_vec(&d_fg, fg, stream));
    RB_TRY(upload_vec(&d_fpg, fpg, stream));
    RB_TRY(upload_vec(&d_fo, fo, stream));
    RB_TRY(upload_vec(&d_frag, frag, stream));
    RB_TRY(upload_vec(&d_ptab, ptab, stream));
    RB_TRY(upload_vec(&d_opts, lay.options, stream));
    const int nopt = a.out_size;
    RB_TRY(hipMalloc(reinterpret_cast<void**>(&d_partial), size_t(nopt) * a.nitems * 2 * sizeof(double)));
    RB_TRY(hipMalloc(reinterpret_cast<void**>(&d_price), nopt * sizeof(double)));
    RB_TRY(hipMalloc(reinterpret_cast<void**>(&d_stat), nopt * sizeof(double)));
    RB_TRY(hipMalloc(reinterpret_cast<void**>(&d_iv), nopt * sizeof(double)));
    a.units = d_fu;
    a.groups = d_fg;
    a.pgs = d_fpg;
    a.opts = d_fo;
    a.frag = d_frag;
    a.ptab = d_ptab;
    a.partial = d_partial;
    nblk = a.nitems;
    chunk = req.M;
    RB_TRY(hipStreamSynchronize(stream));
    return ST_OK;
}

#define RB_LAUNCH(id, call)            \
    do {                               \
        const int slot_ = timed(id);   \
        RB_TRY(call);                  \
        mark(slot_, nullptr, nullptr); \
    } while (0)

int Plan::run_fused(std::string& err) {
    const int nopt = fa.out_size;
    begin_timing();
    RB_LAUNCH(K_FUSED, launch_fused(fused_mode, fused_upw, fa, fused_grid, stream));
    RB_LAUNCH(K_REDUCE, launch_reduce_f64(d_partial, fa.nitems, req.M, nopt, d_price, d_stat, stream));
    RB_LAUNCH(K_IV, launch_iv_f64(d_price, d_opts, nopt, d_iv, stream));
    RB_TRY(hipStreamSynchronize(stream));
    end_timing();
    return ST_OK;
}

}  // namespace rb
