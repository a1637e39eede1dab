// rb_server.cpp -- batched pricing and the continuous-batching server.
//
// The reference prices one request per call, with the process's threads all working on it
// (ComputeIVRTMT, rBergLib/rBergomiMT.cpp:301-425); its clients fan requests out over
// processes (rBergomiR/rbergomi_console.R:150-153).  Here requests are planned on the host as
// they arrive and priced together: a batch is one launch of the fused kernel over the work
// items of all its requests.  The server keeps up to max_inflight batches on the device, each
// on its own stream, and admits pending requests as soon as the device bytes of the batches in
// flight leave room, so new work starts while earlier batches drain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rbergomi_mi355x.h"
#include "rb_engine.h"
#include "rb_fused_host.h"

using namespace rb;
using Clock = std::chrono::steady_clock;

#define RB_TRY_SRV(expr)                                                     \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return e_ == hipErrorOutOfMemory ? ST_ERR_NOMEM : ST_ERR_DEVICE; \
        }                                                                    \
    } while (0)

namespace {

struct Job {
    long ticket = 0;
    Request req;
    Layout lay;
    FusedHostReq fh;
    double* price = nullptr;
    double* iv = nullptr;
    double* stat = nullptr;
    int out = 0;
    size_t footprint = 0;
    Clock::time_point t_submit, t_done;
    int status = -1;  // -1 pending, else an rb_status
    std::string err;  // message of a failed job
    bool done = false;
    bool long_path = false;  // N > 1024: priced alone by the single-launch FFT kernel
    double conv_flops = 0.0;  // algorithmic Toeplitz flops: M N (N - 1) per distinct H
};

// LDS bytes one workgroup may use, and per CU, on the current device
int lds_limits(size_t& block, size_t& cu, std::string& err) {
    int device = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&device) != hipSuccess || device_props(device, prop) != hipSuccess) {
        err = "no HIP device";
        return ST_ERR_DEVICE;
    }
    block = prop.sharedMemPerBlock;
    cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    return ST_OK;
}

// the request's inputs (not its output pointers) as a fused-kernel Request
int request_of(const rb_request* q, int precision, int method, int comp, Request& r, std::string& err) {
    r.xi = q->xi;
    if (q->par_size < 1 || q->op_size < 1 || !q->H || !q->eta || !q->rho || !q->T || !q->K) {
        err = "empty or null parameter arrays";
        return ST_ERR_ARG;
    }
    r.H.assign(q->H, q->H + q->par_size);
    r.eta.assign(q->eta, q->eta + q->par_size);
    r.rho.assign(q->rho, q->rho + q->par_size);
    r.T.assign(q->T, q->T + q->op_size);
    r.K.assign(q->K, q->K + q->op_size);
    r.N = q->N;
    r.M = q->M;
    if (q->seed && q->seed_size > 0) r.seed.assign(q->seed, q->seed + q->seed_size);
    r.method = method;
    r.comp = comp;
    r.precision = precision;
    if (fused_mode_of(precision) < 0) {
        err = "batches run the fused kernel (RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM)";
        return ST_ERR_ARG;
    }
    return validate(r, err);
}

int plan_job(const rb_request* q, int precision, int method, int comp, Job& j, std::string& err) {
    if (!q->price || !q->stat) {
        err = "null output arrays";
        return ST_ERR_ARG;
    }
    int rc = request_of(q, precision, method, comp, j.req, err);
    if (rc) return rc;
    const Request& r = j.req;
    j.lay = make_layout(r);
    j.price = q->price;
    j.iv = q->iv;
    j.stat = q->stat;
    j.out = r.out_size();
    j.conv_flops = double(r.M) * double(r.N) * double(r.N - 1) * double(j.lay.units.size());
    size_t lds_block = 0, lds_cu = 0;
    rc = lds_limits(lds_block, lds_cu, err);
    if (rc) return rc;
    if (r.N > 1024) {
        // long paths: one request per launch of the FFT kernel
        j.long_path = true;
        return FftFused::plan(r, j.lay, lds_block, j.footprint, err);
    }
    rc = plan_fused_request(r, j.lay, fused_mode_of(precision), j.fh, err);
    if (rc) return rc;
    // alone it must fit one workgroup's LDS; batches are closed before a request that would push
    // their combined LDS over the limit
    const size_t lds = fused_batch_lds({&j.fh}, lds_cu, nullptr, nullptr);
    if (lds > lds_block) {
        err = "request needs " + std::to_string(lds) + " bytes of LDS, above the " + std::to_string(lds_block) +
              " of one workgroup (too many options or maturities for its N)";
        return ST_ERR_LIMIT;
    }
    j.footprint = FusedBatch::footprint({&j.fh});
    return ST_OK;
}

struct Flight {
    std::unique_ptr<FusedBatch> batch;
    std::unique_ptr<FftFused> fft;  // a long-path request (then jobs holds just it)
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t kbeg = nullptr, kend = nullptr;  // around the pricing kernel (timing events)
    double* pinned = nullptr;  // [3][total_out]: price, iv, stat
    std::vector<Job*> jobs;
    size_t footprint = 0;
};

void destroy_events(Flight& F) {
    for (hipEvent_t* e : {&F.done, &F.kbeg, &F.kend})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
}

}  // namespace

struct rb_server {
    rb_server_config cfg{};
    int device = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> pending;
    std::map<long, std::unique_ptr<Job>> jobs;
    std::vector<std::unique_ptr<Flight>> flights;
    std::vector<hipStream_t> free_streams;
    std::map<hipStream_t, std::pair<double*, size_t>> pinned;  // reusable result staging per stream
    // reusable upload staging per stream (pinned: the batch's one table copy runs without the
    // server thread waiting for it; a stream is reused only after its flight has completed)
    std::map<hipStream_t, std::pair<unsigned char*, size_t>> upinned;
    size_t inflight_bytes = 0;
    bool stop = false;
    long next_ticket = 1;
    std::vector<double> latencies;
    double path_steps = 0.0;
    long batches = 0;
    int method = METHOD_RT, comp = COMP_REF;
    size_t lds_block = 0, lds_cu = 0;
    // device timing: each flight's pricing kernel as [begin, end] ms after `epoch`
    hipEvent_t epoch = nullptr;
    std::vector<std::pair<double, double>> intervals;
    double kernel_ms = 0.0, conv_flops = 0.0;
    long peak_inflight = 0;
    // timing events around each batch's kernel (rb_server_set_timing), off by default.  The epoch
    // is recorded on a flight's stream with the first timed batch: an event recorded once on the
    // null stream made every later batch's kernel run after the previous one (serving closed loop
    // 1.29e11 against 1.72e11 paths*steps/s on MI355X), while events on the flights' own streams
    // leave their kernels overlapping
    bool timing = false;

    void loop();
    bool reap(std::unique_lock<std::mutex>& lk);
    bool launch(std::unique_lock<std::mutex>& lk);
};

bool rb_server::reap(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    for (size_t f = 0; f < flights.size();) {
        Flight& F = *flights[f];
        const hipError_t q = hipEventQuery(F.done);
        if (q == hipErrorNotReady) {
            ++f;
            continue;
        }
        const int total = F.batch ? F.batch->total_out : F.fft->total_out;
        const Clock::time_point now = Clock::now();
        if (q == hipSuccess && F.kbeg) {
            float a = 0.f, b = 0.f;
            if (epoch && F.kend && hipEventElapsedTime(&a, epoch, F.kbeg) == hipSuccess &&
                hipEventElapsedTime(&b, epoch, F.kend) == hipSuccess) {
                intervals.push_back({double(a), double(b)});
                kernel_ms += double(b) - double(a);
            }
            (void)hipGetLastError();  // a failed query must not be reported by the next launch
        }
        for (size_t i = 0; i < F.jobs.size(); ++i) {
            Job& j = *F.jobs[i];
            if (q == hipSuccess) {
                const long o = F.batch ? F.batch->out_off[i] : 0;
                std::memcpy(j.price, F.pinned + o, j.out * sizeof(double));
                if (j.iv) std::memcpy(j.iv, F.pinned + total + o, j.out * sizeof(double));
                std::memcpy(j.stat, F.pinned + 2 * total + o, j.out * sizeof(double));
                j.status = ST_OK;
                latencies.push_back(std::chrono::duration<double, std::milli>(now - j.t_submit).count());
                path_steps += double(j.req.M) * j.req.N;
                conv_flops += j.conv_flops;
            } else {
                j.status = ST_ERR_DEVICE;
                j.err = std::string("batch execution failed on the device: ") + hipGetErrorString(q);
            }
            j.t_done = now;
            j.done = true;
        }
        inflight_bytes -= F.footprint;
        destroy_events(F);
        F.batch.reset();
        F.fft.reset();
        free_streams.push_back(F.stream);
        flights.erase(flights.begin() + f);
        any = true;
    }
    if (any) cv_done.notify_all();
    (void)lk;
    return any;
}

bool rb_server::launch(std::unique_lock<std::mutex>& lk) {
    if (int(flights.size()) >= cfg.max_inflight || pending.empty()) return false;
    // admit from the front of the queue while the device budget allows
    std::vector<Job*> take;
    std::vector<const FusedHostReq*> tables;
    size_t bytes = 0;
    double work = 0.0;
    while (!pending.empty() && int(take.size()) < cfg.max_batch_requests) {
        Job* j = pending.front();
        // a long-path request runs alone; it closes the batch before it
        if (j->long_path && !take.empty()) break;
        if (inflight_bytes + bytes + j->footprint > size_t(cfg.max_device_bytes)) break;
        const double w = double(j->req.M) * j->req.N;
        if (!take.empty() && cfg.max_batch_path_steps > 0.0 && work + w > cfg.max_batch_path_steps) break;
        if (!j->long_path) {
            // one launch holds the largest LDS terms of its requests: close the batch before a
            // request that would push them over one workgroup's LDS (alone it fits), or out of
            // two workgroups per CU
            tables.push_back(&j->fh);
            if (!take.empty() && !fused_batch_accepts(tables, lds_block, lds_cu)) break;
        }
        take.push_back(j);
        bytes += j->footprint;
        work += w;
        pending.pop_front();
        if (j->long_path) break;
    }
    if (take.empty()) return false;
    inflight_bytes += bytes;
    std::unique_ptr<Flight> F(new Flight());
    F->jobs = take;
    F->footprint = bytes;
    if (!free_streams.empty()) {
        F->stream = free_streams.back();
        free_streams.pop_back();
    } else if (hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking) != hipSuccess) {
        F->stream = nullptr;
    }
    lk.unlock();
    std::string err = F->stream ? "" : "cannot create a stream";
    int rc = F->stream ? ST_OK : ST_ERR_DEVICE;
    if (!rc && timing && (hipEventCreate(&F->kbeg) != hipSuccess || hipEventCreate(&F->kend) != hipSuccess)) {
        err = "cannot create timing events";
        rc = ST_ERR_DEVICE;
    }
    if (!rc && timing && !epoch) {
        // the zero of the device timeline, created with the first timed batch, on the flight's
        // stream (not the null stream, see `timing`)
        if (hipEventCreate(&epoch) != hipSuccess || hipEventRecord(epoch, F->stream) != hipSuccess ||
            hipEventSynchronize(epoch) != hipSuccess) {
            err = "cannot create the timing epoch";
            rc = ST_ERR_DEVICE;
        }
    }
    int total = 0;
    double* d_price = nullptr;
    double* d_iv = nullptr;
    double* d_stat = nullptr;
    if (take[0]->long_path) {
        F->fft.reset(new FftFused());
        if (!rc) rc = F->fft->build(take[0]->req, take[0]->lay, device, F->stream, err);
        if (!rc) rc = F->fft->launch(F->stream, err, nullptr, F->kbeg, F->kend);
        if (!rc) {
            total = F->fft->total_out;
            d_price = F->fft->d_price;
            d_iv = F->fft->d_iv;
            d_stat = F->fft->d_stat;
        }
    } else {
        std::vector<const FusedHostReq*> hr;
        for (Job* j : take) hr.push_back(&j->fh);
        F->batch.reset(new FusedBatch());
        auto& up = upinned[F->stream];  // only the server thread touches the pool
        F->batch->stage = [&up](size_t bytes) -> unsigned char* {
            if (up.second < bytes) {
                if (up.first) (void)hipHostFree(up.first);
                up.first = nullptr;
                up.second = 0;
                if (hipHostMalloc(reinterpret_cast<void**>(&up.first), 2 * bytes) == hipSuccess) up.second = 2 * bytes;
            }
            return up.first;
        };
        const int mode = fused_mode_of(cfg.precision);
        if (!rc) rc = F->batch->build(hr, mode, device, F->stream, err);
        if (!rc) rc = F->batch->launch(F->stream, err, nullptr, F->kbeg, F->kend);
        if (!rc) {
            total = F->batch->total_out;
            d_price = F->batch->d_price;
            d_iv = F->batch->d_iv;
            d_stat = F->batch->d_stat;
        }
    }
    if (!rc) {
        auto& pb = pinned[F->stream];  // only the server thread touches the pool
        const size_t need = size_t(std::max(1, 3 * total));
        if (pb.second < need) {
            if (pb.first) (void)hipHostFree(pb.first);
            pb.first = nullptr;
            pb.second = 0;
            if (hipHostMalloc(reinterpret_cast<void**>(&pb.first), 2 * need * sizeof(double)) == hipSuccess)
                pb.second = 2 * need;
        }
        F->pinned = pb.first;
        if (!F->pinned) rc = ST_ERR_NOMEM;
    }
    if (!rc) {
        const size_t n = size_t(total) * sizeof(double);
        // one copy when the results lie price | iv | stat in device memory (fused batches)
        const bool adjacent = d_iv == d_price + total && d_stat == d_price + 2 * size_t(total);
        if ((adjacent ? hipMemcpyAsync(F->pinned, d_price, 3 * n, hipMemcpyDeviceToHost, F->stream)
                      : hipMemcpyAsync(F->pinned, d_price, n, hipMemcpyDeviceToHost, F->stream) ||
                            hipMemcpyAsync(F->pinned + total, d_iv, n, hipMemcpyDeviceToHost, F->stream) ||
                            hipMemcpyAsync(F->pinned + 2 * total, d_stat, n, hipMemcpyDeviceToHost, F->stream)) ||
            hipEventCreateWithFlags(&F->done, hipEventDisableTiming) ||
            hipEventRecord(F->done, F->stream)) {
            err = "cannot enqueue the result copies";
            rc = ST_ERR_DEVICE;
        }
    }
    lk.lock();
    if (rc) {
        for (Job* j : take) {
            j->status = rc;
            j->err = err;
            j->done = true;
            j->t_done = Clock::now();
        }
        inflight_bytes -= bytes;
        if (F->stream) (void)hipStreamSynchronize(F->stream);
        destroy_events(*F);
        F->batch.reset();
        F->fft.reset();
        if (F->stream) free_streams.push_back(F->stream);
        cv_done.notify_all();
        return true;
    }
    ++batches;
    flights.push_back(std::move(F));
    peak_inflight = std::max(peak_inflight, long(flights.size()));
    return true;
}

void rb_server::loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (stop && pending.empty() && flights.empty()) break;
        bool progress = reap(lk);
        progress = launch(lk) || progress;
        if (!progress) {
            // new requests wake the loop at once; completions are polled
            cv_work.wait_for(lk, std::chrono::microseconds(flights.empty() ? 2000 : 100));
        }
    }
}

// A planned set of requests priced together, run many times (rb_batch_*): the short requests
// (N <= 1024) in as few fused launches as one workgroup's LDS allows, each long one in a launch of
// the FFT kernel, all on one stream; optionally recorded once into a hipGraph and replayed.
struct rb_batch {
    int device = 0;
    rb_config cfg{};
    hipStream_t st = nullptr;
    bool own_stream = false;
    std::vector<std::unique_ptr<Job>> jobs;
    std::vector<int> short_idx, long_idx;
    std::vector<std::pair<int, int>> runs;  // index ranges into short_idx, one fused launch each
    std::vector<std::unique_ptr<FusedBatch>> fused;
    std::vector<std::unique_ptr<FftFused>> longs;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;

    ~rb_batch() {
        drop_graph();
        fused.clear();
        longs.clear();
        if (st) (void)hipStreamSynchronize(st);
        if (own_stream && st) (void)hipStreamDestroy(st);
    }
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    int plan(const rb_request* reqs, int n, std::vector<std::unique_ptr<Job>>& out, std::string& err) const {
        out.clear();
        for (int i = 0; i < n; ++i) {
            out.emplace_back(new Job());
            const int rc = plan_job(&reqs[i], cfg.precision, cfg.method, cfg.compensator, *out.back(), err);
            if (rc) {
                err = "request " + std::to_string(i) + ": " + err;
                return rc;
            }
        }
        return ST_OK;
    }
    int build(std::string& err) {
        short_idx.clear();
        long_idx.clear();
        std::vector<const FusedHostReq*> hr;
        for (int i = 0; i < int(jobs.size()); ++i) {
            if (jobs[i]->long_path) {
                long_idx.push_back(i);
            } else {
                short_idx.push_back(i);
                hr.push_back(&jobs[i]->fh);
            }
        }
        size_t lds_block = 0, lds_cu = 0;
        int rc = lds_limits(lds_block, lds_cu, err);
        if (rc) return rc;
        runs = FusedBatch::split(hr, lds_block, lds_cu, false);
        for (const std::pair<int, int>& run : runs) {
            fused.emplace_back(new FusedBatch());
            const std::vector<const FusedHostReq*> part(hr.begin() + run.first, hr.begin() + run.second);
            rc = fused.back()->build(part, fused_mode_of(cfg.precision), device, st, err);
            if (rc) return rc;
        }
        for (int i : long_idx) {
            longs.emplace_back(new FftFused());
            rc = longs.back()->build(jobs[i]->req, jobs[i]->lay, device, st, err);
            if (rc) return rc;
        }
        return ST_OK;
    }
    int enqueue(std::string& err) {
        for (auto& b : fused) {
            const int rc = b->launch(st, err);
            if (rc) return rc;
        }
        for (auto& f : longs) {
            const int rc = f->launch(st, err);
            if (rc) return rc;
        }
        return ST_OK;
    }
    int run(std::string& err);
};

int rb_batch::run(std::string& err) {
    if (use_graph && !gexec) {
        // one eager run sets kernel attributes outside the capture; the graph then records the
        // same launches (kernels only: the counter is cleared by a kernel, results are copied
        // out after the replay)
        int rc = enqueue(err);
        if (rc) return rc;
        RB_TRY_SRV(hipStreamSynchronize(st));
        RB_TRY_SRV(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = enqueue(err);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        RB_TRY_SRV(e);
        graph = g;
        RB_TRY_SRV(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    if (use_graph) {
        RB_TRY_SRV(hipGraphLaunch(gexec, st));
    } else {
        const int rc = enqueue(err);
        if (rc) return rc;
    }
    for (size_t r = 0; r < runs.size(); ++r)
        for (int k = runs[r].first; k < runs[r].second; ++k) {
            const Job& j = *jobs[short_idx[k]];
            const int rc = fused[r]->results(k - runs[r].first, j.price, j.iv, j.stat, st, err);
            if (rc) return rc;
        }
    for (size_t l = 0; l < long_idx.size(); ++l) {
        const Job& j = *jobs[long_idx[l]];
        const int rc = longs[l]->results(j.price, j.iv, j.stat, st, err);
        if (rc) return rc;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "batch execution failed on the device";
        return ST_ERR_DEVICE;
    }
    return ST_OK;
}

extern "C" {

int rb_batch_create(const rb_request* reqs, int n, const rb_config* cfg, void* stream, rb_batch** out) {
    set_last_error("");
    if (!reqs || n < 1 || !cfg || !out) {
        set_last_error("need at least one request, a configuration and an output handle");
        return RB_ERR_ARG;
    }
    std::unique_ptr<rb_batch> b(new rb_batch());
    b->cfg = *cfg;
    std::string err;
    int rc = b->plan(reqs, n, b->jobs, err);
    if (!rc && hipGetDevice(&b->device) != hipSuccess) {
        err = "no HIP device";
        rc = RB_ERR_DEVICE;
    }
    if (!rc) {
        if (stream) {
            b->st = static_cast<hipStream_t>(stream);
        } else if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess) {
            b->own_stream = true;
        } else {
            err = "cannot create a stream";
            rc = RB_ERR_DEVICE;
        }
    }
    if (!rc) rc = b->build(err);
    if (rc) {
        set_last_error(err);
        return rc;
    }
    *out = b.release();
    return RB_OK;
}

int rb_batch_update(rb_batch* b, const rb_request* reqs, int n) {
    set_last_error("");
    if (!b || !reqs || n != int(b->jobs.size())) {
        set_last_error("the update must give as many requests as the batch holds");
        return RB_ERR_ARG;
    }
    std::string err;
    std::vector<std::unique_ptr<Job>> nj;
    int rc = b->plan(reqs, n, nj, err);
    for (int i = 0; !rc && i < n; ++i)
        if (nj[i]->long_path != b->jobs[i]->long_path || nj[i]->req.N != b->jobs[i]->req.N ||
            nj[i]->req.M != b->jobs[i]->req.M || nj[i]->out != b->jobs[i]->out) {
            err = "request " + std::to_string(i) + ": N, M and the option count must stay as they were";
            rc = RB_ERR_ARG;
        }
    if (rc) {
        set_last_error(err);
        return rc;
    }
    // short requests: the new tables into the same device buffers (a captured graph stays valid);
    // every launch's layout is checked before any is changed
    std::vector<const FusedHostReq*> hr;
    for (int i : b->short_idx) hr.push_back(&nj[i]->fh);
    for (int pass = 0; pass < 2; ++pass)
        for (size_t r = 0; !rc && r < b->runs.size(); ++r) {
            const std::vector<const FusedHostReq*> part(hr.begin() + b->runs[r].first,
                                                        hr.begin() + b->runs[r].second);
            rc = b->fused[r]->build(part, fused_mode_of(b->cfg.precision), b->device, b->st, err, true, pass == 0);
        }
    // long requests carry their Philox key in the kernel arguments: rebuilt, and the graph is
    // recorded again on the next run
    for (size_t l = 0; !rc && l < b->long_idx.size(); ++l) {
        const Job& j = *nj[b->long_idx[l]];
        rc = b->longs[l]->build(j.req, j.lay, b->device, b->st, err);
        if (!rc) b->drop_graph();
    }
    if (rc) {
        set_last_error(err);
        return rc;
    }
    b->jobs.swap(nj);
    return RB_OK;
}

int rb_batch_set_graph(rb_batch* b, int enable) {
    if (!b) return RB_ERR_ARG;
    b->use_graph = enable != 0;
    if (!b->use_graph) b->drop_graph();
    return RB_OK;
}

int rb_batch_run(rb_batch* b) {
    set_last_error("");
    if (!b) return RB_ERR_ARG;
    std::string err;
    const int rc = b->run(err);
    if (rc) set_last_error(err);
    return rc;
}

void rb_batch_destroy(rb_batch* b) { delete b; }

int rb_price_batch(const rb_request* reqs, int n, const rb_config* cfg) {
    rb_batch* b = nullptr;
    int rc = rb_batch_create(reqs, n, cfg, nullptr, &b);
    if (!rc) rc = rb_batch_run(b);
    rb_batch_destroy(b);
    return rc;
}

int rb_autotune_fused(const rb_request* shape, const int* N, int nN, const rb_config* cfg, int* pb_out,
                      double* ms_out) {
    set_last_error("");
    if (!shape || !N || nN < 1 || !cfg) {
        set_last_error("need a request shape, at least one N and a configuration");
        return RB_ERR_ARG;
    }
    Request r;
    std::string err;
    rb_request q = *shape;
    q.N = N[0];
    int rc = request_of(&q, cfg->precision, cfg->method, cfg->compensator, r, err);
    if (rc) {
        set_last_error(err);
        return rc;
    }
    for (int i = 0; i < nN; ++i)
        if (N[i] < 1 || N[i] > 1024) {
            set_last_error("the fused kernel covers 1 <= N <= 1024");
            return RB_ERR_LIMIT;
        }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        set_last_error("no HIP device");
        return RB_ERR_DEVICE;
    }
    std::vector<int> chosen;
    std::vector<double> ms;
    rc = autotune_fused(r, std::vector<int>(N, N + nN), fused_mode_of(cfg->precision), device, chosen, ms, err);
    if (rc) {
        set_last_error(err);
        return rc;
    }
    for (int i = 0; i < nN; ++i) {
        if (pb_out) pb_out[i] = chosen[i];
        if (ms_out) ms_out[i] = ms[i];
    }
    return RB_OK;
}

int rb_tune_save(const char* path) {
    set_last_error("");
    std::string err;
    const int rc = path ? tune_save(path, err) : RB_ERR_ARG;
    if (rc) set_last_error(path ? err : "null path");
    return rc;
}

int rb_tune_load(const char* path) {
    set_last_error("");
    std::string err;
    const int rc = path ? tune_load(path, err) : RB_ERR_ARG;
    if (rc) set_last_error(path ? err : "null path");
    return rc;
}

int rb_tune_clear(void) {
    tune_clear();
    return RB_OK;
}

int rb_tune_table(int* rows, int max) {
    const std::vector<std::array<int, 4>> t = tune_table();
    for (int i = 0; rows && i < max && i < int(t.size()); ++i)
        for (int c = 0; c < 4; ++c) rows[4 * i + c] = t[i][c];
    return int(t.size());
}

int rb_server_create(const rb_server_config* cfg, rb_server** out) {
    return rb_server_create_method(cfg, RB_METHOD_RT, RB_COMP_REF, out);
}

int rb_server_create_method(const rb_server_config* cfg, int method, int compensator, rb_server** out) {
    set_last_error("");
    if (!cfg || !out) {
        set_last_error("null argument");
        return RB_ERR_ARG;
    }
    if ((method != RB_METHOD_RT && method != RB_METHOD_MC) || (compensator != RB_COMP_REF && compensator != RB_COMP_EXACT)) {
        set_last_error("unknown estimator or compensator");
        return RB_ERR_ARG;
    }
    if (fused_mode_of(cfg->precision) < 0) {
        set_last_error("the server runs the fused kernel (RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM)");
        return RB_ERR_ARG;
    }
    std::unique_ptr<rb_server> s(new rb_server());
    s->cfg = *cfg;
    s->method = method;
    s->comp = compensator;
    {
        std::string err;
        if (lds_limits(s->lds_block, s->lds_cu, err) != ST_OK) {
            set_last_error(err);
            return RB_ERR_DEVICE;
        }
    }
    s->cfg.max_inflight = std::max(1, cfg->max_inflight);
    s->cfg.max_batch_requests = std::max(1, cfg->max_batch_requests);
    if (s->cfg.max_device_bytes <= 0) s->cfg.max_device_bytes = 8L << 30;
    if (hipGetDevice(&s->device) != hipSuccess) {
        set_last_error("no HIP device");
        return RB_ERR_DEVICE;
    }
    rb_server* p = s.get();
    s->th = std::thread([p] { p->loop(); });
    *out = s.release();
    return RB_OK;
}

int rb_server_submit(rb_server* s, const rb_request* req, long* ticket) {
    if (!s || !req || !ticket) return RB_ERR_ARG;
    std::unique_ptr<Job> j(new Job());
    j->t_submit = Clock::now();
    std::string err;
    const int rc = plan_job(req, s->cfg.precision, s->method, s->comp, *j, err);
    if (rc) {
        set_last_error(err);
        return rc;
    }
    if (j->footprint > size_t(s->cfg.max_device_bytes)) {
        set_last_error("request needs " + std::to_string(j->footprint) + " device bytes, above the server budget");
        return RB_ERR_LIMIT;
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        j->ticket = s->next_ticket++;
        *ticket = j->ticket;
        s->pending.push_back(j.get());
        s->jobs[j->ticket] = std::move(j);
    }
    s->cv_work.notify_one();
    return RB_OK;
}

int rb_server_wait(rb_server* s, long ticket, double* latency_ms) {
    if (!s) return RB_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) {
        set_last_error("unknown ticket");
        return RB_ERR_ARG;
    }
    Job* j = it->second.get();
    s->cv_done.wait(lk, [j] { return j->done; });
    const int st = j->status;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    if (st != ST_OK) set_last_error(j->err);
    s->jobs.erase(it);
    return st;
}

int rb_server_stats(rb_server* s, double* p50_ms, double* p99_ms, long* completed, double* path_steps,
                    long* batches) {
    if (!s) return RB_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {  // nearest rank
        if (l.empty()) return 0.0;
        const long k = long(std::ceil(p * double(l.size()))) - 1;
        return l[size_t(std::max(0L, std::min(long(l.size()) - 1, k)))];
    };
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (completed) *completed = long(l.size());
    if (path_steps) *path_steps = s->path_steps;
    if (batches) *batches = s->batches;
    return RB_OK;
}

int rb_server_set_timing(rb_server* s, int enable) {
    if (!s) return RB_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    s->timing = enable != 0;
    return RB_OK;
}

int rb_server_report_get(rb_server* s, rb_server_report* out) {
    if (!s || !out) return RB_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    rb_server_report r{};
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {  // nearest rank
        if (l.empty()) return 0.0;
        const long k = long(std::ceil(p * double(l.size()))) - 1;
        return l[size_t(std::max(0L, std::min(long(l.size()) - 1, k)))];
    };
    r.p50_ms = pct(0.50);
    r.p99_ms = pct(0.99);
    r.completed = long(l.size());
    r.path_steps = s->path_steps;
    r.batches = s->batches;
    r.peak_inflight = s->peak_inflight;
    r.kernel_launches = long(s->intervals.size());
    r.kernel_ms = s->kernel_ms;
    r.conv_flops = s->conv_flops;
    // sweep over the kernels' device intervals: union, span and the most overlapping at once
    std::vector<std::pair<double, int>> ev;
    for (const auto& iv : s->intervals) {
        ev.push_back({iv.first, +1});
        ev.push_back({iv.second, -1});
    }
    std::sort(ev.begin(), ev.end());  // at equal times an end (-1) sorts before a begin
    int live = 0;
    double last = 0.0;
    for (const auto& e : ev) {
        if (live > 0) r.device_busy_ms += e.first - last;
        if (live > 1) r.device_overlap_ms += e.first - last;
        live += e.second;
        r.max_concurrent = std::max(r.max_concurrent, long(live));
        last = e.first;
    }
    if (!ev.empty()) r.device_span_ms = ev.back().first - ev.front().first;
    *out = r;
    return RB_OK;
}

void rb_server_destroy(rb_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    for (hipStream_t st : s->free_streams) (void)hipStreamSynchronize(st);
    for (auto& kv : s->pinned)
        if (kv.second.first) (void)hipHostFree(kv.second.first);
    for (auto& kv : s->upinned)
        if (kv.second.first) (void)hipHostFree(kv.second.first);
    for (hipStream_t st : s->free_streams) (void)hipStreamDestroy(st);
    if (s->epoch) (void)hipEventDestroy(s->epoch);
    delete s;
}

}  // extern "C"
