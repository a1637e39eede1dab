"""Benchmark of the rough-Bergomi Romano-Touzi pricing path on one MI355X.

Workload (one "step" = one full pricing call over device-resident work buffers):
  single parameter set at scale -- H = 0.1, eta = 1.5, rho = -0.7, xi = 0.04, N = 512 steps,
  M = 2^22 paths, maturities {0.1, 0.25, 0.5, 1, 2} x 41 log-moneyness points in [-0.4, 0.2]
  (205 options), Philox Gaussians generated on the device, FP64 price accumulation.
Metric: paths*steps/s = M*N / seconds per pricing call.

Prints one JSON line (rank 0, N = 1 GPU only) with the roofline of the dominant kernel, timed
live with HIP events on the plan's stream, and a CPU baseline: the CPU restatement of the
reference's OpenMP driver (mt19937_64 Gaussians, per-path FFT convolution, options loop)
timed on a bounded sample of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MATURITIES = [0.1, 0.25, 0.5, 1.0, 2.0]
LOG_MONEYNESS = np.linspace(-0.4, 0.2, 41)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0}
PREC_CODE = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3}


def workload(paths):
    T = np.repeat(MATURITIES, len(LOG_MONEYNESS))
    K = np.exp(np.tile(LOG_MONEYNESS, len(MATURITIES)))
    return dict(xi=0.04, H=[0.1], eta=[1.5], rho=[-0.7], T=T, K=K, N=512, M=paths)


def load_traffic(kernel):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    ent = data.get(kernel)
    return None if ent is None else ent.get("bytes_per_launch")


def cpu_baseline(sample_paths, threads):
    from oracle import pyoracle as orc
    w = workload(sample_paths)
    t0 = time.perf_counter()
    orc.compute_rt(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"], [2024, 1],
                   num_threads=threads, conv=orc.CONV_FFT)
    dt = time.perf_counter() - t0
    return {"value": sample_paths * w["N"] / dt, "unit": "paths*steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d paths of the same workload (N=512, 205 options), %d OpenMP threads, "
                      "%.1f s" % (sample_paths, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp16", choices=sorted(PREC_CODE))
    ap.add_argument("--paths", type=int, default=2 ** 22)
    ap.add_argument("--cpu-sample-paths", type=int, default=2 ** 18)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run the streaming copy probe over this many MiB (traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(paths would shard by range with one all-reduce of the option sums)", file=sys.stderr)
        sys.exit(2)

    import rbergomi_amd as rb

    w = workload(args.paths)
    plan = rb.Plan(w["xi"], w["H"], w["eta"], w["rho"], w["T"], w["K"], w["N"], w["M"],
                   [20240601, 7], precision=PREC_CODE[args.precision])
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
                   [20240601, 7], precision=PREC_CODE[args.precision], profile=True)
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
    if "volterra" in dominant or "fused" in dominant:
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS[args.precision]
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": load_traffic(dominant), "kernel": dominant,
                "avg_launch_ms": avg_ms, "launches_per_step": launches,
                "flops_per_launch": flops_per_launch}
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
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16x2+f32", "bf16": "bf16x2+f32"}[args.precision],
        "data": "synthetic (device Philox4x32-10 Gaussians)",
        "config": {"workload": "single parameter set at scale: H=0.1 eta=1.5 rho=-0.7 xi=0.04 "
                               "N=512 M=2^22, 5 maturities x 41 strikes",
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
                      [20240601, 7], precision=PREC_CODE["fp64"])
        ref.run()
        p64, iv64, s64 = ref.results()
        ref.close()
        out["accuracy"] = {"max_abs_price_err_vs_fp64": float(np.max(np.abs(price - p64))),
                           "max_abs_stat_err_vs_fp64": float(np.max(np.abs(stat - s64))),
                           "max_price_stat_ratio": float(np.max(np.abs(price - p64) / s64)),
                           "reference": "staged FP64 path, same Philox counters"}
    if args.copy_probe:
        import ctypes
        ms_copy = ctypes.c_double()
        nbytes = args.copy_probe << 20
        assert rb.load().rb_probe_copy(nbytes, 5, ctypes.byref(ms_copy)) == 0
        out["copy_probe"] = {"bytes": nbytes, "avg_ms": ms_copy.value,
                             "GB_per_s": 2 * nbytes / (ms_copy.value * 1e-3) / 1e9}
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_paths, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
