// rb_fused_host.h -- host-side tables and batches of the fused mixed-precision kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
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

int plan_fused_request(const Request& req, const Layout& lay, int mode, FusedHostReq& out,
                       std::string& err);

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
    double* d_partial = nullptr;
    double* d_price = nullptr;
    double* d_stat = nullptr;
    double* d_iv = nullptr;
    Option* d_ivopts = nullptr;
    long* d_row_off = nullptr;
    long* d_row_n = nullptr;
    long* d_row_M = nullptr;

    ~FusedBatch();
    void release();
    // device bytes the batch would allocate
    static size_t footprint(const std::vector<const FusedHostReq*>& reqs);
    int build(const std::vector<const FusedHostReq*>& reqs, int mode, int device, hipStream_t st,
              std::string& err);
    // enqueue: zero the item counter, fused kernel, ordered reduction, implied vols
    int launch(hipStream_t st, std::string& err, Plan* timing = nullptr);
    // enqueue the copies of request q's results (caller synchronizes the stream)
    int results(int q, double* price, double* iv, double* stat, hipStream_t st, std::string& err);
};

}  // namespace rb
