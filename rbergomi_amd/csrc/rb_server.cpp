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
    bool done = false;
    bool long_path = false;  // N > 1024: priced alone by the single-launch FFT kernel
};

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
    if (r.N > 1024) {
        // long paths: one request per launch of the FFT kernel
        int device = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&device) != hipSuccess || device_props(device, prop) != hipSuccess) {
            err = "no HIP device";
            return ST_ERR_DEVICE;
        }
        j.long_path = true;
        return FftFused::plan(r, j.lay, prop.sharedMemPerBlock, j.footprint, err);
    }
    rc = plan_fused_request(r, j.lay, fused_mode_of(precision), j.fh, err);
    if (rc) return rc;
    j.footprint = FusedBatch::footprint({&j.fh});
    return ST_OK;
}

struct Flight {
    std::unique_ptr<FusedBatch> batch;
    std::unique_ptr<FftFused> fft;  // a long-path request (then jobs holds just it)
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    double* pinned = nullptr;  // [3][total_out]: price, iv, stat
    std::vector<Job*> jobs;
    size_t footprint = 0;
};

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
    size_t inflight_bytes = 0;
    bool stop = false;
    long next_ticket = 1;
    std::vector<double> latencies;
    double path_steps = 0.0;
    long batches = 0;
    std::string fatal;

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
            } else {
                j.status = ST_ERR_DEVICE;
            }
            j.t_done = now;
            j.done = true;
        }
        inflight_bytes -= F.footprint;
        (void)hipEventDestroy(F.done);
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
    size_t bytes = 0;
    double work = 0.0;
    while (!pending.empty() && int(take.size()) < cfg.max_batch_requests) {
        Job* j = pending.front();
        // a long-path request runs alone; it closes the batch before it
        if (j->long_path && !take.empty()) break;
        if (inflight_bytes + bytes + j->footprint > size_t(cfg.max_device_bytes)) break;
        const double w = double(j->req.M) * j->req.N;
        if (!take.empty() && cfg.max_batch_path_steps > 0.0 && work + w > cfg.max_batch_path_steps) break;
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
    std::string err;
    int rc = F->stream ? ST_OK : ST_ERR_DEVICE;
    int total = 0;
    double* d_price = nullptr;
    double* d_iv = nullptr;
    double* d_stat = nullptr;
    if (take[0]->long_path) {
        F->fft.reset(new FftFused());
        if (!rc) rc = F->fft->build(take[0]->req, take[0]->lay, device, F->stream, err);
        if (!rc) rc = F->fft->launch(F->stream, err);
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
        const int mode = fused_mode_of(cfg.precision);
        if (!rc) rc = F->batch->build(hr, mode, device, F->stream, err);
        if (!rc) rc = F->batch->launch(F->stream, err);
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
        if (hipMemcpyAsync(F->pinned, d_price, n, hipMemcpyDeviceToHost, F->stream) ||
            hipMemcpyAsync(F->pinned + total, d_iv, n, hipMemcpyDeviceToHost, F->stream) ||
            hipMemcpyAsync(F->pinned + 2 * total, d_stat, n, hipMemcpyDeviceToHost, F->stream) ||
            hipEventCreateWithFlags(&F->done, hipEventDisableTiming) ||
            hipEventRecord(F->done, F->stream))
            rc = ST_ERR_DEVICE;
    }
    lk.lock();
    if (rc) {
        for (Job* j : take) {
            j->status = rc;
            j->done = true;
            j->t_done = Clock::now();
        }
        inflight_bytes -= bytes;
        if (F->stream) (void)hipStreamSynchronize(F->stream);
        if (F->done) (void)hipEventDestroy(F->done);
        F->batch.reset();
        F->fft.reset();
        if (F->stream) free_streams.push_back(F->stream);
        fatal = err;
        cv_done.notify_all();
        return true;
    }
    ++batches;
    flights.push_back(std::move(F));
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

extern "C" {

int rb_price_batch(const rb_request* reqs, int n, const rb_config* cfg) {
    set_last_error("");
    if (!reqs || n < 1 || !cfg) {
        set_last_error("need at least one request and a configuration");
        return RB_ERR_ARG;
    }
    std::vector<std::unique_ptr<Job>> jobs;
    std::string err;
    for (int i = 0; i < n; ++i) {
        jobs.emplace_back(new Job());
        const int rc = plan_job(&reqs[i], cfg->precision, cfg->method, cfg->compensator, *jobs.back(), err);
        if (rc) {
            set_last_error("request " + std::to_string(i) + ": " + err);
            return rc;
        }
    }
    int device = 0;
    hipStream_t st = nullptr;
    if (hipGetDevice(&device) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        set_last_error("no HIP device");
        return RB_ERR_DEVICE;
    }
    // the requests up to N = 1024 in one launch, each longer one in a launch of its own
    std::vector<const FusedHostReq*> hr;
    std::vector<int> short_idx;
    for (int i = 0; i < n; ++i)
        if (!jobs[i]->long_path) {
            hr.push_back(&jobs[i]->fh);
            short_idx.push_back(i);
        }
    FusedBatch B;
    int rc = RB_OK;
    if (!hr.empty()) {
        rc = B.build(hr, fused_mode_of(cfg->precision), device, st, err);
        if (!rc) rc = B.launch(st, err);
        for (size_t k = 0; !rc && k < short_idx.size(); ++k) {
            const rb_request& q = reqs[short_idx[k]];
            rc = B.results(int(k), q.price, q.iv, q.stat, st, err);
        }
    }
    for (int i = 0; !rc && i < n; ++i) {
        if (!jobs[i]->long_path) continue;
        FftFused F;
        rc = F.build(jobs[i]->req, jobs[i]->lay, device, st, err);
        if (!rc) rc = F.launch(st, err);
        if (!rc) rc = F.results(reqs[i].price, reqs[i].iv, reqs[i].stat, st, err);
        if (!rc && hipStreamSynchronize(st) != hipSuccess) {
            err = "long-path execution failed";
            rc = RB_ERR_DEVICE;
        }
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) {
        err = "batch execution failed";
        rc = RB_ERR_DEVICE;
    }
    B.release();
    (void)hipStreamDestroy(st);
    if (rc) set_last_error(err);
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
    set_last_error("");
    if (!cfg || !out) {
        set_last_error("null argument");
        return RB_ERR_ARG;
    }
    if (fused_mode_of(cfg->precision) < 0) {
        set_last_error("the server runs the fused kernel (RB_PREC_SPLIT_FP16, RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM)");
        return RB_ERR_ARG;
    }
    std::unique_ptr<rb_server> s(new rb_server());
    s->cfg = *cfg;
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
    const int rc = plan_job(req, s->cfg.precision, METHOD_RT, COMP_REF, *j, err);
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
    if (st != ST_OK) set_last_error(s->fatal);
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
    for (hipStream_t st : s->free_streams) (void)hipStreamDestroy(st);
    delete s;
}

}  // extern "C"
