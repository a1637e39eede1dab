"""Benchmark of the rough-Bergomi Romano-Touzi pricing path on one MI355X.

Default line -- the served mixed-precision pricer on the serving trace:
  4096 calibration requests, each a parameter set (H ~ U[0.02, 0.45], eta ~ U[0.5, 4],
  rho ~ U[-0.99, -0.1], xi ~ U[0.01, 0.16]) with 10-200 options over 1-5 maturities
  T ~ U[0.02, 2] and log-moneyness U[-0.5, 0.3], a path budget M = 2^U, U ~ U[14, 21], and
  N = max(32, ceil(256 T_max)) steps, priced by the continuous-batching server on the split-FP16
  Toeplitz-GEMM kernel with Philox Gaussians generated on the device and FP64 price sums.
  One "step" = the whole trace submitted at once and every result waited for (closed loop).
Metric: paths*steps/s = sum of M*N over the trace / seconds per step, with the p50 / p99 request
latency of the same trace under Poisson arrivals at 0.7 x the measured capacity, and the largest
absolute price error against the FP64 path on the same Philox counters.

The line carries the roofline of the dominant kernel (the fused pricing kernel of every batch,
timed with HIP events on each batch's stream), the single-set workload (N = 512, M = 2^22,
205 options) as a sub-object with its own roofline, and a CPU baseline: the CPU restatement of the
reference's OpenMP driver (mt19937_64 Gaussians, per-path FFT convolution, options loop) timed on
a bounded sample of the same trace on the host cores, with all the threads the host gives the job
and with one thread.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MATURITIES = [0.1, 0.25, 0.5, 1.0, 2.0]
LOG_MONEYNESS = np.linspace(-0.4, 0.2, 41)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0, "fp32gemm": 157.3}
# fp32: FP32 FFT convolution (staged); fp32gemm: Toeplitz GEMM on FP32 MFMA (fused kernel);
# fp32fft: FP32 FFT convolution inside the single-launch kernel
PREC_CODE = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3, "fp32gemm": 4, "fp32fft": 5}


def workload(paths):
    T = np.repeat(MATURITIES, len(LOG_MONEYNESS))
    K = np.exp(np.tile(LOG_MONEYNESS, len(MATURITIES)))
    return dict(xi=0.04, H=[0.1], eta=[1.5], rho=[-0.7], T=T, K=K, N=512, M=paths)


def valu_roofline(rb, prof, avg_ms):
    """The vector-issue side of the fused kernel's roofline.  The launch's instructions by class
    (rocprofv3 --pmc counts of one launch, committed in profiles/traffic.json) priced at the issue
    costs this device shows for each class (rb_probe_issue, measured now): FP32 FMA-class, FP32
    transcendental, 64-bit integer (priced as the 32 x 32 -> 64-bit multiply-add, the Philox
    product), FP64.  Matrix instructions run on the matrix cores and are left out.  The sum over
    the device's SIMDs, as a fraction of the launch, is the share of the launch that one SIMD needs
    just to issue its vector instructions."""
    ns = (ctypes.c_double * 4)()
    simds = ctypes.c_int()
    assert rb.load().rb_probe_issue(ns, ctypes.byref(simds)) == 0
    fma, trans, madpair, f64 = list(ns)
    mad = max(madpair - fma, fma)          # the probe's pair is the multiply-add and a 32-bit add
    v = prof["valu_insts_per_launch"]
    tr = prof["trans_f32_insts_per_launch"]
    i64 = prof.get("int64_insts_per_launch", 0.0)
    d64 = prof.get("f64_insts_per_launch", 0.0)
    mf = prof.get("mfma_insts_per_launch", 0.0)
    plain = v - tr - i64 - d64 - mf
    busy_ns = (plain * fma + tr * trans + i64 * mad + d64 * f64) / simds.value
    return {"insts_per_launch": v,
            "by_class": {"fp32": plain, "trans_f32": tr, "int64": i64, "f64": d64, "mfma": mf},
            "ns_per_inst": {"fp32": fma, "trans_f32": trans, "int64": mad, "f64": f64},
            "simds": simds.value, "issue_busy_ms": busy_ns * 1e-6, "issue_busy_frac": busy_ns / (avg_ms * 1e6),
            "source": "SQ_INSTS_VALU, _TRANS_F32, _INT64, _FMA/ADD/MUL/TRANS_F64, SQ_INSTS_MFMA (rocprofv3 --pmc, %s); "
                      "issue costs from rb_probe_issue on this device" % prof.get("source")}


def load_profile(kernel):
    """The committed rocprofv3 --pmc summary of the dominant kernel on the default workload:
    HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    VALU / transcendental instruction counts per launch, or {}."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    return data.get(kernel, {})


def host_cpu():
    """The host's CPU model and thread counts (lscpu, os)."""
    model = ""
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """The host threads this job may use: OMP_NUM_THREADS where the host sets it (the GPU box
    sets the share of one GPU's job), else every core of the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = len(os.sched_getaffinity(0))
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_time_requests(reqs, threads, m_scale):
    """The oracle's restatement of the reference's OpenMP driver (compute_rt, FFT convolution)
    over requests with M scaled by m_scale; returns (paths*steps, seconds)."""
    from oracle import pyoracle as orc
    work = 0
    t0 = time.perf_counter()
    for d in reqs:
        M = max(1, int(round(d["M"] * m_scale)))
        orc.compute_rt(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], M, d["seed"],
                       num_threads=threads, conv=orc.CONV_FFT)
        work += M * d["N"]
    return work, time.perf_counter() - t0


def cpu_baseline_serving(reqs, threads, n_multi=512, n_single=32, m_scale=1.0 / 256):
    """CPU baseline of the serving trace: the first n_multi requests with every M scaled by
    m_scale on `threads` threads, and the first n_single of them on one thread."""
    w, dt = cpu_time_requests(reqs[:n_multi], threads, m_scale)
    w1, dt1 = cpu_time_requests(reqs[:n_single], 1, m_scale)
    return {"value": w / dt, "unit": "paths*steps/s", "cores": threads, "kind": "port",
            "sample": "first %d requests of the serving trace with M scaled by 1/%d (%.3g paths*steps), "
                      "oracle restatement of the reference's OpenMP driver (mt19937_64, per-path FFT "
                      "convolution, options loop), %d OpenMP threads, %.1f s"
                      % (n_multi, round(1 / m_scale), w, threads, dt),
            "single_thread": {"value": w1 / dt1, "sample": "first %d requests, M scaled by 1/%d, %.1f s"
                                                           % (n_single, round(1 / m_scale), dt1)},
            "host": host_cpu()}


def cpu_baseline(sample_paths, threads, method="rt"):
    from oracle import pyoracle as orc
    w = workload(sample_paths)
    fn = orc.compute_rt if method == "rt" else orc.compute_mc
    t0 = time.perf_counter()
    fn(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"], [2024, 1],
       num_threads=threads, conv=orc.CONV_FFT)
    dt = time.perf_counter() - t0
    return {"value": sample_paths * w["N"] / dt, "unit": "paths*steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d paths of the same workload (N=512, 205 options, %s estimator), "
                      "%d OpenMP threads, %.1f s" % (sample_paths, method.upper(), threads, dt)}


def batched_requests(nsets=256, seed=3):
    """256 parameter sets drawn from a seed: H~U[0.02,0.45], eta~U[0.5,4], rho~U[-0.99,-0.1],
    xi~U[0.01,0.16], T~U[0.05,2], N uniform on {32,64,128,256,512}, M = 2^U with U~U[14,20],
    21 strikes evenly in [0.8, 1.2] at T."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsets):
        H, eta, rho = rng.uniform(0.02, 0.45), rng.uniform(0.5, 4.0), rng.uniform(-0.99, -0.1)
        xi, T = rng.uniform(0.01, 0.16), rng.uniform(0.05, 2.0)
        N = int(rng.choice([32, 64, 128, 256, 512]))
        M = int(round(2 ** rng.uniform(14, 20)))
        out.append(dict(xi=xi, H=[H], eta=[eta], rho=[rho], T=[T] * 21, K=np.linspace(0.8, 1.2, 21),
                        N=N, M=M, seed=[seed, i]))
    return out


def serving_requests(n, seed=5):
    """Calibration requests: parameters as in the batched workload, 10-200 options spread over
    1-5 maturities T~U[0.02,2] with log-moneyness U[-0.5,0.3], M = 2^U with U~U[14,21],
    N = max(32, ceil(256 T_max))."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        H, eta, rho = rng.uniform(0.02, 0.45), rng.uniform(0.5, 4.0), rng.uniform(-0.99, -0.1)
        xi = rng.uniform(0.01, 0.16)
        mats = rng.uniform(0.02, 2.0, int(rng.integers(1, 6)))
        nopt = int(rng.integers(10, 201))
        T = np.sort(rng.choice(mats, nopt))
        K = np.exp(rng.uniform(-0.5, 0.3, nopt))
        N = max(32, int(np.ceil(256 * mats.max())))
        M = int(round(2 ** rng.uniform(14, 21)))
        out.append(dict(xi=xi, H=[H], eta=[eta], rho=[rho], T=T, K=K, N=N, M=M, seed=[seed, i]))
    return out


# server settings of the latency measurements: four batches in flight, each capped at 2.5e8
# paths*steps (about 1.7 ms of work), chosen by a sweep at fixed offered load
SERVE_INFLIGHT, SERVE_MAX_BATCH, SERVE_CAP = 4, 64, 2.5e8


def run_serving(rb, reqs, prec, rate_fraction=0.7, max_inflight=SERVE_INFLIGHT,
                max_batch=SERVE_MAX_BATCH, cap=SERVE_CAP, rate=None):
    """Closed loop (all submitted at once) gives the capacity; then Poisson arrivals at
    rate_fraction x capacity (or at `rate` requests/s) give the latency percentiles.  cap: the
    server's work cap per batch (sum of M*N, 0 = none)."""
    import threading
    from rbergomi_amd.serving import Request, Server

    def mk(d):
        return Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"])

    work = sum(d["M"] * d["N"] for d in reqs)
    with Server(prec, max_inflight=max_inflight, max_batch_requests=max_batch,
                max_batch_path_steps=cap) as srv:
        rs = [mk(d) for d in reqs]
        t0 = time.perf_counter()
        for r in rs:
            srv.submit(r)
        for r in rs:
            srv.wait(r)
        closed = time.perf_counter() - t0
    capacity = len(reqs) / closed
    if rate is None:
        rate = rate_fraction * capacity
    rng = np.random.default_rng(11)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    with Server(prec, max_inflight=max_inflight, max_batch_requests=max_batch,
                max_batch_path_steps=cap) as srv:
        rs = [mk(d) for d in reqs]
        done = []

        def waiter():
            for r in rs:
                while r.ticket is None:
                    time.sleep(1e-4)
                srv.wait(r)
                done.append(r.latency_ms)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for r, g in zip(rs, gaps):
            t_next += g
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            srv.submit(r)
        th.join()
        open_s = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array(done)
    return {"requests": len(reqs), "paths_steps": float(work),
            "closed_loop_s": closed, "closed_loop_paths_steps_per_s": work / closed,
            "capacity_req_per_s": capacity, "offered_req_per_s": rate,
            "open_loop_s": open_s, "open_loop_paths_steps_per_s": work / open_s,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "server_p50_ms": st["p50_ms"], "server_p99_ms": st["p99_ms"], "batches": st["batches"],
            "server": {"max_inflight": max_inflight, "max_batch_requests": max_batch,
                       "max_batch_path_steps": cap}}


def tune(rb, Ns, args, precisions):
    """Tile-shape autotuning for the step counts of a workload (a 21-strike request shape at
    M = 2^20); returns what was chosen and the time it took."""
    from rbergomi_amd.serving import Request, autotune
    shape = Request(0.04, [0.1], [1.5], [-0.7], [1.0] * 21, list(np.linspace(0.8, 1.2, 21)), 0, 1 << 20, [5, 5])
    t0 = time.perf_counter()
    chosen = {}
    for name, code in precisions:
        res = autotune(shape, sorted(set((n + 31) // 32 * 32 for n in Ns)), precision=code)
        chosen[name] = {str(n): pb for n, (pb, _) in res.items()}
    return {"path_blocks_per_tile": chosen, "tuning_s": time.perf_counter() - t0,
            "shape": "21 strikes at T = 1, M = 2^20, per distinct N"}


def bench_batched(rb, args):
    from rbergomi_amd.serving import Request, price_batch
    reqs = batched_requests()
    work = sum(d["M"] * d["N"] for d in reqs)
    out = {"metric": "paths*steps/s", "unit": "paths*steps/s", "n_gpus": 1, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "data": "synthetic (device Philox4x32-10 Gaussians)",
           "config": {"workload": "batched: 256 parameter sets, N in {32..512}, M = 2^14..2^20, 21 strikes",
                      "paths_steps": work, "parallelism": "1 GPU"}}
    # plain FP64 single-request GPU path: the requests one after another, two passes (the first
    # also pays one-time costs such as first allocations of each size); the faster one counts
    passes = []
    for _ in range(2):
        t0 = time.perf_counter()
        ref = [rb.price_v2(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"],
                           precision=rb.PREC_FP64) for d in reqs]
        passes.append(time.perf_counter() - t0)
    fp64_s = min(passes)
    out["fp64_single_request"] = {"s": fp64_s, "paths_steps_per_s": work / fp64_s, "passes_s": passes}
    precs = (("fp16", rb.PREC_SPLIT_FP16), ("bf16", rb.PREC_SPLIT_BF16), ("fp32gemm", rb.PREC_FP32_GEMM))
    if args.autotune:
        out["autotune"] = tune(rb, [d["N"] for d in reqs], args, precs)
    for name, code in (("fp16", rb.PREC_SPLIT_FP16), ("bf16", rb.PREC_SPLIT_BF16), ("fp32gemm", rb.PREC_FP32_GEMM)):
        rs = [Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"])
              for d in reqs]
        price_batch(rs, precision=code)  # warm-up (code objects, allocator)
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            price_batch(rs, precision=code)
            times.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(times))
        err = [np.max(np.abs(r.price - p[0])) for r, p in zip(rs, ref)]
        ratio = [np.max(np.abs(r.price - p[0]) / np.maximum(p[2], 1e-300)) for r, p in zip(rs, ref)]
        out[name] = {"ms_per_batch": ms, "paths_steps_per_s": work / (ms * 1e-3),
                     "max_abs_price_err_vs_fp64": float(np.max(err)),
                     "median_abs_price_err_vs_fp64": float(np.median(err)),
                     "max_err_over_stat": float(np.max(ratio)),
                     "speedup_vs_fp64_single_request": fp64_s / (ms * 1e-3)}
    out["value"] = out["fp16"]["paths_steps_per_s"]
    out["ms_per_step"] = out["fp16"]["ms_per_batch"]
    out["dtype"] = "f16+f32"
    out["steps"], out["warmup"] = args.steps, 1
    return out


def poisson_latency(rb, ds, rate, seed=11):
    """Poisson arrivals of the requests at `rate` per second through a fresh server with the
    benchmark's settings: latency percentiles (submit to results copied) and throughput."""
    import threading
    from rbergomi_amd.serving import Request, Server
    reqs = [Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"]) for d in ds]
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(reqs))
    done = []
    with Server(rb.PREC_SPLIT_FP16, max_inflight=SERVE_INFLIGHT, max_batch_requests=SERVE_MAX_BATCH,
                max_batch_path_steps=SERVE_CAP) as srv:
        def waiter():
            for r in reqs:
                while r.ticket is None:
                    time.sleep(1e-4)
                srv.wait(r)
                done.append(r.latency_ms)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for r, g in zip(reqs, gaps):
            t_next += g
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            srv.submit(r)
        th.join()
        wall = time.perf_counter() - t0
        rep = srv.report()
    lat = np.array(done)
    work = float(sum(d["M"] * d["N"] for d in ds))
    return {"requests": len(ds), "offered_req_per_s": rate, "wall_s": wall,
            "paths_steps_per_s": work / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "batches": rep["batches"],
            "peak_batches_in_flight": rep["peak_inflight"]}


def serving_line(rb, args):
    """The default line: the serving trace through the continuous-batching server, closed loop;
    then its latency under Poisson arrivals and its price error against the FP64 path."""
    from rbergomi_amd.serving import Request, Server
    ds = serving_requests(args.requests)
    work = float(sum(d["M"] * d["N"] for d in ds))
    reqs = [Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"]) for d in ds]
    with Server(rb.PREC_SPLIT_FP16, max_inflight=SERVE_INFLIGHT, max_batch_requests=SERVE_MAX_BATCH,
                max_batch_path_steps=SERVE_CAP) as srv:
        def step():
            for r in reqs:
                srv.submit(r)
            for r in reqs:  # every result copied back: the device work of the step has finished
                srv.wait(r)

        for _ in range(args.warmup):
            step()
        r0 = srv.report()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        t1 = time.perf_counter()
        r1 = srv.report()
    s_step = (t1 - t0) / args.steps
    assert abs((r1["path_steps"] - r0["path_steps"]) - work * args.steps) <= 1e-9 * work * args.steps
    out = {"metric": "paths*steps/s", "value": work / s_step, "unit": "paths*steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": s_step * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f16+f32",
           "data": "synthetic (seeded parameter trace, device Philox4x32-10 Gaussians)",
           "config": {"workload": "serving trace: %d calibration requests (10-200 options over 1-5 maturities, "
                                  "M = 2^14..2^21, N = max(32, 256 T)), continuous batching, closed loop"
                                  % len(ds),
                      "requests": len(ds), "paths_steps_per_step": work, "precision": "fp16",
                      "server": {"max_inflight": SERVE_INFLIGHT, "max_batch_requests": SERVE_MAX_BATCH,
                                 "max_batch_path_steps": SERVE_CAP},
                      "parallelism": "1 GPU"},
           "batches_per_step": (r1["batches"] - r0["batches"]) / args.steps,
           "peak_batches_in_flight": r1["peak_inflight"]}
    if not args.no_latency:
        lat = poisson_latency(rb, ds, 0.7 * len(ds) / s_step)
        out["latency"] = lat
        out["p50_ms"], out["p99_ms"] = lat["p50_ms"], lat["p99_ms"]
    if not args.no_check:
        # the served results against the FP64 path on the same Philox counters (every 8th request;
        # each request's results are the same alone and in any mix)
        # the same loop times the plain FP64 single-request GPU path (rb_price_v2, one request
        # after another) on these requests: the path the served pricer is compared with
        errs = []
        sub = list(zip(ds, reqs))[::8]
        t64 = time.perf_counter()
        for dd, r in sub:
            p64 = rb.price_v2(dd["xi"], dd["H"], dd["eta"], dd["rho"], dd["T"], dd["K"], dd["N"], dd["M"], dd["seed"])[0]
            errs.append(float(np.max(np.abs(r.price - p64))))
        t64 = time.perf_counter() - t64
        w64 = float(sum(dd["M"] * dd["N"] for dd, _ in sub))
        out["accuracy"] = {"max_abs_price_err_vs_fp64": max(errs), "median_abs_price_err_vs_fp64": float(np.median(errs)),
                           "requests_checked": len(errs), "reference": "FP64 path, same Philox counters"}
        out["fp64_single_request"] = {"requests": len(sub), "s": t64, "paths_steps_per_s": w64 / t64,
                                      "served_speedup": (work / s_step) / (w64 / t64),
                                      "what": "every 8th request of the trace through rb_price_v2 (FP64), one "
                                              "after another, including its host-side planning and copies"}
    return out


def serving_roofline(rb, args):
    """The roofline of the serving trace's dominant kernel: one closed-loop pass of the trace
    through a server with one batch in flight and HIP timing events around every batch's kernel, on
    its stream, so that each kernel's duration is its own (in the timed steps up to four batches'
    kernels share the device, and a kernel's start-to-end time includes its neighbours' work)."""
    from rbergomi_amd.serving import Request, Server
    ds = serving_requests(args.requests)
    work = float(sum(d["M"] * d["N"] for d in ds))
    reqs = [Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"]) for d in ds]
    with Server(rb.PREC_SPLIT_FP16, max_inflight=1, max_batch_requests=SERVE_MAX_BATCH,
                max_batch_path_steps=SERVE_CAP) as srv:
        srv.set_timing(True)
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            srv.wait(r)
        d = srv.report()
    avg_ms = d["kernel_ms"] / d["kernel_launches"]
    flops_launch = d["conv_flops"] / d["kernel_launches"]
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12
    prof = load_profile("k_fused_serving")
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["fp16"], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS["fp16"], "traffic": prof.get("bytes_per_launch"),
            "kernel": "k_fused (the fused pricing kernel of every batch)", "avg_launch_ms": avg_ms,
            "launches_per_step": d["kernel_launches"], "flops_per_launch": flops_launch,
            "timing": "HIP events on each batch's stream around its kernel, in one closed-loop pass of the "
                      "trace with one batch in flight, after the timed steps (which run without events)",
            "timing_pass": {"kernel_ms": d["kernel_ms"], "device_busy_ms": d["device_busy_ms"],
                            "batches": d["batches"],
                            "achieved_over_busy_TFLOPs": d["conv_flops"] / (d["device_busy_ms"] * 1e-3) / 1e12}}
    if prof.get("valu_insts_per_launch"):
        roof["valu"] = valu_roofline(rb, prof, avg_ms)
        roof["valu"]["lane_insts_per_path_step"] = 64.0 * prof["valu_insts_per_launch"] * d["kernel_launches"] / work
    return roof


def bench_serving(rb, args):
    reqs = serving_requests(args.requests)
    tuned = tune(rb, [d["N"] for d in reqs], args, (("fp16", rb.PREC_SPLIT_FP16),)) if args.autotune else None
    res = run_serving(rb, reqs, rb.PREC_SPLIT_FP16)
    out = {"metric": "paths*steps/s", "value": res["open_loop_paths_steps_per_s"], "unit": "paths*steps/s",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f16+f32", "data": "synthetic (device Philox4x32-10 Gaussians)",
           "config": {"workload": "serving: %d calibration requests, Poisson arrivals at 0.7 x capacity, "
                                  "10-200 options over 1-5 maturities, M = 2^14..2^21, N = max(32, 256 T)"
                                  % args.requests, "parallelism": "1 GPU"},
           "serving": res, "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"]}
    if tuned:
        out["autotune"] = tuned
    return out


def bench_config1(rb, args):
    """H = 0.07, eta = 1.9, rho = -0.9, xi = 0.235^2, T = 1, N = 64, M = 2^16, 21 strikes in
    [0.8, 1.2]: each precision through a plan (device-resident buffers), and the CPU restatement
    of the reference's driver on the whole workload (host threads and one thread)."""
    xi, H, eta, rho = 0.235 ** 2, [0.07], [1.9], [-0.9]
    T, K = [1.0] * 21, list(np.linspace(0.8, 1.2, 21))
    N, M = 64, 2 ** 16
    rows = {}
    p64 = None
    for name, code in (("fp64", rb.PREC_FP64), ("fp16", rb.PREC_SPLIT_FP16), ("bf16", rb.PREC_SPLIT_BF16),
                       ("fp32gemm", rb.PREC_FP32_GEMM)):
        pl = rb.Plan(xi, H, eta, rho, T, K, N, M, [20240601, 1], precision=code)
        for _ in range(args.warmup):
            pl.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            pl.run()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        p, _, st = pl.results()
        pl.close()
        rows[name] = {"ms": ms, "paths_steps_per_s": M * N / (ms * 1e-3)}
        if p64 is None:
            p64 = p
            rows[name]["price_atm"] = float(p[10])
            rows[name]["stat_atm"] = float(st[10])
        else:
            rows[name]["max_abs_err_vs_fp64"] = float(np.max(np.abs(p - p64)))
    out = {"metric": "paths*steps/s", "unit": "paths*steps/s", "n_gpus": 1, "higher_is_better": True,
           "value": rows["fp16"]["paths_steps_per_s"], "dtype": "f16+f32", "vs_baseline": None,
           "scaling": "weak", "data": "synthetic (device Philox4x32-10 Gaussians)",
           "config": {"workload": "one parameter set: N=64, M=2^16, 21 strikes"}, "rows": rows}
    if not args.no_cpu:
        from oracle import pyoracle as orc
        cpu = {}
        for th in (args.cpu_threads or cpu_threads(), 1):
            t0 = time.perf_counter()
            orc.compute_rt(xi, H, eta, rho, T, K, N, M, [2024, 1], num_threads=th, conv=orc.CONV_FFT)
            dt = time.perf_counter() - t0
            cpu["%d_threads" % th] = {"s": dt, "paths_steps_per_s": M * N / dt}
        out["cpu_baseline"] = {"whole workload": cpu, "kind": "port", "host": host_cpu()}
    return out


def bench_long(rb, args):
    """N in {1024, 2048, 4096}, H in {0.02, 0.05}, eta 2.3, rho -0.9, xi 0.03, T 1, M = 2^18,
    31 strikes in [0.7, 1.3]: the FP64 path, the FP32 LDS-FFT convolution, and the split-FP16 and
    FP32 Toeplitz GEMMs where they apply (N <= 1024)."""
    K = np.linspace(0.7, 1.3, 31)
    rows = []
    for N in (1024, 2048, 4096):
        for H in (0.02, 0.05):
            row = {"N": N, "H": H}
            # fp16: the split-FP16 GEMM kernel up to N = 1024, the single-launch mixed-precision
            # FFT kernel beyond; fp32fft: the staged pipeline with the FP32 FFT convolution
            for name, code in (("fp64", rb.PREC_FP64), ("fp16", rb.PREC_SPLIT_FP16),
                               ("fp32gemm", rb.PREC_FP32_GEMM), ("fp32fft", rb.PREC_FP32)):
                if name == "fp32gemm" and N > 1024:
                    continue
                pl = rb.Plan(0.03, [H], [2.3], [-0.9], [1.0] * 31, K, N, 2 ** 18, [4, N], precision=code)
                pl.run()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    pl.run()
                ms = (time.perf_counter() - t0) * 1e3 / args.steps
                p, _, st = pl.results()
                pl.close()
                row[name] = {"ms": ms, "paths_steps_per_s": 2 ** 18 * N / (ms * 1e-3)}
                row[name + "_price_atm"] = float(p[15])
                if name != "fp64":
                    row[name + "_max_abs_err_vs_fp64"] = float(np.max(np.abs(p - row["_p64"])))
                else:
                    row["_p64"] = p
            row.pop("_p64", None)
            rows.append(row)
    cpu = None
    if not args.no_cpu:
        # the CPU restatement of the reference's driver on 2^12 paths of each (N, H = 0.05)
        from oracle import pyoracle as orc
        cpu = {"kind": "port", "host": host_cpu(), "threads": args.cpu_threads or cpu_threads(), "rows": {}}
        for N in (1024, 2048, 4096):
            t0 = time.perf_counter()
            orc.compute_rt(0.03, [0.05], [2.3], [-0.9], [1.0] * 31, list(K), N, 2 ** 12, [4, N],
                           num_threads=cpu["threads"], conv=orc.CONV_FFT)
            dt = time.perf_counter() - t0
            cpu["rows"][str(N)] = {"paths": 2 ** 12, "s": dt, "paths_steps_per_s": 2 ** 12 * N / dt}
    return {"metric": "paths*steps/s", "unit": "paths*steps/s", "n_gpus": 1, "higher_is_better": True,
            "cpu_baseline": cpu,
            "value": max(r["fp32fft"]["paths_steps_per_s"] for r in rows), "dtype": "f32 FFT convolution, f64 rest",
            "config": {"workload": "long paths: N in {1024,2048,4096}, M = 2^18, 31 strikes"},
            "rows": rows, "vs_baseline": None, "scaling": "weak", "data": "synthetic"}


def single_line(rb, args, sub=False):
    """The single-set workload (N = 512, M = 2^22, 205 options) through a plan; sub: as a
    sub-object of the default line (no latency sample, no CPU baseline)."""
    w = workload(args.paths)
    method = rb.METHOD_RT if args.method == "rt" else rb.METHOD_MC
    plan = rb.Plan(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"],
                   [20240601, 7], precision=PREC_CODE[args.precision], method=method)
    for _ in range(args.warmup):
        plan.run()
    # rb_plan_run returns after synchronizing the stream it launched on
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    price, iv, stat = plan.results()
    plan.close()

    # one more run with HIP events around every launch, on the plan's own stream
    prof = rb.Plan(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"],
                   [20240601, 7], precision=PREC_CODE[args.precision], method=method, profile=True)
    prof.run()
    times = prof.kernel_times()
    info = prof.info()
    prof.close()
    dominant = max(times, key=lambda k: times[k][0])
    total_ms, launches = times[dominant]
    avg_ms = total_ms / launches
    N, M = w["N"], w["M"]
    # algorithmic flops of the Toeplitz convolution: M*N*(N-1) per pricing call
    flops_per_launch = (M / launches) * N * (N - 1)
    roof = None
    # the matrix-core roofline applies to the Toeplitz products; the FFT kernels run on the VALU
    if ("volterra" in dominant or "fused" in dominant) and "fft" not in dominant:
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS[args.precision]
        default = (args.paths == 2 ** 22 and args.precision == "fp16" and args.method == "rt")
        prof_counts = load_profile(dominant) if "fused" in dominant and default else {}
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": prof_counts.get("bytes_per_launch"),
                "kernel": dominant, "avg_launch_ms": avg_ms, "launches_per_step": launches,
                "flops_per_launch": flops_per_launch}
        if prof_counts.get("valu_insts_per_launch"):
            roof["valu"] = valu_roofline(rb, prof_counts, avg_ms)
    out = {
        "metric": "paths*steps/s",
        "value": M * N / (ms * 1e-3),
        "unit": "paths*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16+f32", "bf16": "bf16x2+f32",
                  "fp32gemm": "f32", "fp32fft": "f32"}[args.precision],
        "data": "synthetic (device Philox4x32-10 Gaussians)",
        "config": {"workload": "single parameter set at scale: H=0.1 eta=1.5 rho=-0.7 xi=0.04 "
                               "N=512 M=2^22, 5 maturities x 41 strikes"
                               + ("" if args.method == "rt" else ", plain Monte Carlo estimator"),
                   "paths": M, "steps": N, "options": len(w["T"]), "precision": args.precision,
                   "parallelism": "1 GPU"},
        "kernel_ms": {k: v[0] for k, v in times.items()},
        "plan": info,
        "roofline": roof,
        "check": {"price_atm_T1": float(price[3 * 41 + 27]), "stat_atm_T1": float(stat[3 * 41 + 27])},
    }
    if not args.no_check and args.precision != "fp64":
        # same Philox counters through the FP64 path: the price error of the mixed precision
        ref = rb.Plan(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"],
                      [20240601, 7], precision=PREC_CODE["fp64"], method=method)
        ref.run()
        p64, iv64, s64 = ref.results()
        ref.close()
        out["accuracy"] = {"max_abs_price_err_vs_fp64": float(np.max(np.abs(price - p64))),
                           "max_abs_stat_err_vs_fp64": float(np.max(np.abs(stat - s64))),
                           "max_price_stat_ratio": float(np.max(np.abs(price - p64) / s64)),
                           "reference": "staged FP64 path, same Philox counters"}
    if not sub and not args.no_latency and args.precision in ("fp16", "bf16", "fp32gemm") and args.method == "rt":
        # request latency under load: 256 serving requests, Poisson arrivals at 0.7 x capacity
        lat = run_serving(rb, serving_requests(256), PREC_CODE[args.precision])
        out["latency"] = {k: lat[k] for k in ("requests", "p50_ms", "p99_ms", "offered_req_per_s",
                                               "open_loop_paths_steps_per_s", "batches")}
        out["p99_ms"] = lat["p99_ms"]
    if args.copy_probe:
        out["copy_probe"] = copy_probe(rb, args.copy_probe)
    if not sub and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_paths, args.cpu_threads or cpu_threads(), args.method)
    return out




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp16", choices=sorted(PREC_CODE))
    ap.add_argument("--workload", default="trace", choices=["trace", "single", "config1", "batched", "serving", "long"],
                    help="trace: the serving trace, closed loop (the default line); single: the single-set "
                         "workload; batched: 256 sets in one launch; serving: Poisson latency sweep; long: "
                         "long paths")
    ap.add_argument("--method", default="rt", choices=["rt", "mc"],
                    help="single workload: Romano-Touzi (computeIVRT_tensor) or plain Monte Carlo "
                         "(computeIV_tensor)")
    ap.add_argument("--requests", type=int, default=4096, help="serving trace size")
    ap.add_argument("--paths", type=int, default=2 ** 22)
    ap.add_argument("--single-steps", type=int, default=10, help="trace: timed steps of the single-set sub-line")
    ap.add_argument("--cpu-sample-paths", type=int, default=2 ** 18)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="trace: skip the single-set sub-line")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 comparison")
    ap.add_argument("--no-latency", action="store_true", help="skip the latency sample")
    ap.add_argument("--autotune", action="store_true",
                    help="batched / serving workloads: autotune the fused kernel's tile shapes first")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run the streaming copy probe over this many MiB (traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(paths would shard by range with one all-reduce of the option sums)", file=sys.stderr)
        sys.exit(2)

    import rbergomi_amd as rb

    if args.workload == "trace":
        out = serving_line(rb, args)
        if not args.no_single:
            sa = argparse.Namespace(**vars(args))
            sa.steps, sa.warmup, sa.precision, sa.method, sa.paths = args.single_steps, 2, "fp16", "rt", 2 ** 22
            out["single_set"] = single_line(rb, sa, sub=True)
        out["roofline"] = serving_roofline(rb, args)
        if args.copy_probe:
            out["copy_probe"] = copy_probe(rb, args.copy_probe)
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_serving(serving_requests(args.requests), args.cpu_threads or cpu_threads())
        print(json.dumps(out), flush=True)
        return
    if args.workload != "single":
        fn = {"batched": bench_batched, "serving": bench_serving, "long": bench_long,
              "config1": bench_config1}[args.workload]
        print(json.dumps(fn(rb, args)), flush=True)
        return
    print(json.dumps(single_line(rb, args)), flush=True)


def copy_probe(rb, mib):
    ms_copy = ctypes.c_double()
    nbytes = mib << 20
    assert rb.load().rb_probe_copy(nbytes, 5, ctypes.byref(ms_copy)) == 0
    return {"bytes": nbytes, "avg_ms": ms_copy.value, "GB_per_s": 2 * nbytes / (ms_copy.value * 1e-3) / 1e9}


if __name__ == "__main__":
    main()
