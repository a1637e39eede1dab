// rb_fused_host.h -- host-side tables and batches of the fused mixed-precision kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <array>
#include <functional>
#include <vector>

#include "rb_fused.h"
#include "rb_types.h"

namespace rb {

struct Request;
struct Layout;
struct Plan;

// One request's tables before upload (offsets are request-local).
struct FusedHostReq {
    FusedReq r{};
    int upw = 1;
    int gmax = 1;
    double item_cost = 0.0;
    std::vector<FusedUnit> fu;
    std::vector<FusedGroup> fg;
    std::vector<FusedPG> fpg;
    std::vector<FusedOption> fo;
    std::vector<uint4> frag;
    std::vector<float> ptab;
    std::vector<Option> ivopts;  // per output I: K, T for the implied vols
};

// pb: path blocks of 32 per tile (1, 2 or 4); 0 takes fused_pb_choice
int plan_fused_request(const Request& req, const Layout& lay, int mode, FusedHostReq& out,
                       std::string& err, int pb = 0);

// Units, groups, payoff groups, options and P_j tables (Np entries per unit) of a request
void fused_tables(const Request& req, const Layout& lay, int Np, int frag_stride, int pshift, FusedHostReq& out,
                  int& pgmax, int& gmax);

// A request on the single-launch FFT kernel (Romano-Touzi or plain MC): long paths (N > 1024)
// at the mixed precisions, any N <= 8192 at RB_PREC_FP32_FFT.
struct FftFused {
    FftFusedArgs args{};
    int grid = 0, total_out = 0;
    hipStream_t stream = nullptr;
    FusedUnit* d_fu = nullptr;
    FusedGroup* d_fg = nullptr;
    FusedPG* d_fpg = nullptr;
    FusedOption* d_fo = nullptr;
    float* d_ptab = nullptr;
    float2* d_ghat = nullptr;
    float2* d_tw = nullptr;
    double* d_partial = nullptr;
    int* d_counter = nullptr;
    long* d_row_off = nullptr;
    long* d_row_n = nullptr;
    long* d_row_M = nullptr;
    double* d_price = nullptr;
    double* d_stat = nullptr;
    double* d_iv = nullptr;
    Option* d_ivopts = nullptr;

    ~FftFused();
    void release();
    // device bytes the request would allocate, after checking N and the LDS it needs (lds_limit)
    static int plan(const Request& req, const Layout& lay, size_t lds_limit, size_t& footprint,
                    std::string& err);
    int build(const Request& req, const Layout& lay, int device, hipStream_t st, std::string& err);
    int launch(hipStream_t st, std::string& err, Plan* timing = nullptr, hipEvent_t kbeg = nullptr,
               hipEvent_t kend = nullptr);
    int results(double* price, double* iv, double* stat, hipStream_t st, std::string& err);
};

// Tile shape (path blocks per tile) per (padded N, operand mode, plain MC): an autotuned entry,
// else the measured default rule.  The table is process-wide and thread-safe; the environment
// variable RBERGOMI_TUNE_CACHE names a file (tune_save's format) loaded on first use.
int fused_default_pb(int NB);
int fused_pb_choice(int Np, int mode, int mc);
void tune_set(int Np, int mode, int mc, int pb);
void tune_clear();
int tune_save(const char* path, std::string& err);
int tune_load(const char* path, std::string& err);
std::vector<std::array<int, 4>> tune_table();
// Times every valid tile shape for each N with `shape`'s parameters, options and M on `device`
// and records the fastest; chosen[i] / best_ms[i] report it for Ns[i].
int autotune_fused(const Request& shape, const std::vector<int>& Ns, int mode, int device,
                   std::vector<int>& chosen, std::vector<double>& best_ms, std::string& err);

// Device memory handed to a batch by its caller: buffers are carved from [base, base + size) at
// 256-byte alignment instead of being allocated, and nothing is freed on release.
struct DevArena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

// LDS of one fused launch over `reqs`: the largest W1 tile, cross-wave partials, per-path records
// and option sums of its requests.  When the partials are what keeps the launch above half of a
// CU's LDS (lds_cu / 2), their chunk of groups is cut so that two workgroups share the CU (the
// same sums in the same order, so the same bits); *cap receives that chunk cap (0: none).  The
// LDS fields of *a are set; returns the bytes one workgroup needs, the kernel's static request
// record included.
size_t fused_batch_lds(const std::vector<const FusedHostReq*>& reqs, size_t lds_cu, FusedBatchArgs* a, int* cap);

// Whether the last request of `reqs` may join the launch of the others: the launch must fit one
// workgroup's LDS (lds_block), and a launch that fits two workgroups per CU (half of lds_cu) is
// not pushed above that, since its requests then run on half the waves.  (A request refused here
// goes to another launch, which the server runs alongside on its own stream.)
bool fused_batch_accepts(const std::vector<const FusedHostReq*>& reqs, size_t lds_block, size_t lds_cu);

// Any number of requests priced by one launch of the fused kernel.
struct FusedBatch {
    int mode = 0, upw = 1, grid = 0, nreq = 0, total_out = 0;
    bool mc = false;  // plain Monte Carlo kernel
    int greg = FUSED_GREG;  // groups per pass of the kernel
    hipStream_t stream = nullptr;  // the stream the buffers were allocated on (and are freed on)
    FusedBatchArgs args{};
    std::vector<long> out_off;   // first output of request q in the result arrays
    FusedReq* d_reqs = nullptr;
    FusedUnit* d_fu = nullptr;
    FusedGroup* d_fg = nullptr;
    FusedPG* d_fpg = nullptr;
    FusedOption* d_fo = nullptr;
    uint4* d_frag = nullptr;
    float* d_ptab = nullptr;
    int* d_item_req = nullptr;
    int* d_item_local = nullptr;
    int* d_counter = nullptr;
    bool counter_dirty = true;   // the item counter may be nonzero (zeroed before the next launch)
    double* d_partial = nullptr;
    double* d_price = nullptr;
    double* d_stat = nullptr;
    double* d_iv = nullptr;
    Option* d_ivopts = nullptr;
    long* d_row_off = nullptr;
    long* d_row_n = nullptr;
    long* d_row_M = nullptr;
    char* d_block = nullptr;     // the one device allocation every pointer above points into

    DevArena* arena = nullptr;   // when set, buffers come from it (see DevArena)
    // pinned host memory of at least the bytes asked for, or null: when set, build packs the
    // tables there, copies them in one piece and returns without waiting for the copy (the
    // caller keeps the memory untouched until the batch's stream has passed the launch);
    // otherwise each table is copied from its host vector and build waits for the copies
    std::function<unsigned char*(size_t)> stage;

    ~FusedBatch();
    void release();
    // device bytes the batch would allocate
    static size_t footprint(const std::vector<const FusedHostReq*>& reqs);
    std::vector<size_t> shape_;  // table sizes and launch shape of the last build (see refill)
    // refill: the batch exists and reqs have the same shapes; their tables are copied into the
    // existing device buffers (pointers, launch shape and work items unchanged).  check_only: only
    // whether a refill with reqs would succeed (nothing changes)
    int build(const std::vector<const FusedHostReq*>& reqs, int mode, int device, hipStream_t st,
              std::string& err, bool refill = false, bool check_only = false);
    // LDS-aware split of reqs into consecutive runs that each fit one launch on a device with
    // lds_block bytes per workgroup and lds_cu per CU: [begin, end) index ranges
    // (soft: also keep launches that fit two workgroups per CU there, fused_batch_accepts)
    static std::vector<std::pair<int, int>> split(const std::vector<const FusedHostReq*>& reqs, size_t lds_block,
                                                  size_t lds_cu, bool soft);
    // enqueue: fused kernel, ordered reduction with implied vols (which zeroes the item counter
    // for the next launch; the counter is zeroed first only after a failed launch); kbeg / kend
    // (may be null) are recorded on st right before and after the fused kernel
    int launch(hipStream_t st, std::string& err, Plan* timing = nullptr, hipEvent_t kbeg = nullptr,
               hipEvent_t kend = nullptr);
    // enqueue the copies of request q's results (caller synchronizes the stream)
    int results(int q, double* price, double* iv, double* stat, hipStream_t st, std::string& err);
};

}  // namespace rb
