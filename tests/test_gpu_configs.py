"""GPU parity on the benchmark configurations themselves, against the CPU oracle.

The mixed-precision kernels (k_fused: split FP16 / BF16 and FP32-operand Toeplitz GEMMs;
k_fused_fft: the single-launch LDS-FFT kernel) are checked directly against the oracle on the
same Philox counters, not only against the FP64 HIP path.  Semantics follow
ComputePayoffRTsamples / ComputeIVRTMT (rBergLib/rBergomiMT.cpp:301-425, 528-610) as restated by
the oracle; the tolerance form follows testRBergLib/tests.cpp:16-35.

Tolerances: FP64 path <= 1e-12 absolute in price and standard error against the oracle on the
same Gaussians; mixed precisions <= 1e-4 absolute per price (the stated mixed-precision budget).

Configurations (synthetic inputs, seeds fixed here):
  1  H=0.07 eta=1.9 rho=-0.9 xi=0.235^2 T=1 N=64 M=2^16, 21 strikes in [0.8, 1.2]
  3  256 parameter sets of bench.batched_requests() (uneven N and M) in one batched launch
  4  N in {1024, 2048, 4096}, H in {0.02, 0.05}, eta=2.3 rho=-0.9 xi=0.03, 31 strikes in [0.7, 1.3]
     (M reduced so that the oracle finishes in seconds)
  5  a 384-request slice of bench.serving_requests() through the continuous-batching server
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CODES = {"fp16": "PREC_SPLIT_FP16", "bf16": "PREC_SPLIT_BF16", "fp32gemm": "PREC_FP32_GEMM"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


@pytest.fixture(scope="module")
def bench():
    import bench as b
    return b


C1 = dict(xi=0.235 ** 2, H=[0.07], eta=[1.9], rho=[-0.9], T=[1.0] * 21,
          K=list(np.linspace(0.8, 1.2, 21)), N=64, M=2 ** 16)
C1_SEED = [20240601, 1]


@pytest.fixture(scope="module")
def n64_oracle(orc):
    c = C1
    Z = orc.normals_philox(C1_SEED, 0, c["M"], 2, c["N"])
    pay = orc.payoff_rt_samples(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], Z)
    return orc.price_stat(pay)


def test_n64_fp64_legacy_abi_matches_oracle(rb, orc, n64_oracle):
    # computeIVRT_tensor itself (the drop-in symbol), num_threads = 8 as a hint
    c = C1
    say("N=64 M=2^16, 21 strikes: computeIVRT_tensor")
    res = rb.computeIV_RT(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], c["M"], C1_SEED, 8)
    pr, st = n64_oracle
    assert np.max(np.abs(res["price"] - pr)) <= 1e-12
    assert np.max(np.abs(res["stat"] - st)) <= 1e-12
    for I in range(21):
        want = orc.iv_call(pr[I], 1.0, c["K"][I], 1.0)
        assert abs(res["iv"][I] - want) <= 1e-9
    # the same prices through the status-returning entry point
    p2, _, s2 = rb.price_v2(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], c["M"], C1_SEED)
    assert np.array_equal(p2, res["price"]) and np.array_equal(s2, res["stat"])


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32gemm"])
def test_n64_mixed_plans_match_oracle(rb, orc, n64_oracle, prec):
    c = C1
    say("N=64 M=2^16: %s plan vs oracle" % prec)
    pl = rb.Plan(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], c["M"], C1_SEED,
                 precision=getattr(rb, CODES[prec]))
    pl.run()
    p, iv, s = pl.results()
    pl.close()
    pr, st = n64_oracle
    err = np.max(np.abs(p - pr))
    say("   max |price - oracle| = %.3e, max |stat - oracle| = %.3e" % (err, np.max(np.abs(s - st))))
    assert err <= 1e-4
    assert np.max(np.abs(s - st)) <= 1e-4
    # the prices are close to the reference's own accuracy bar as well: 0.005 + stat of
    # testRBergLib/tests.cpp:34 is far looser; here within 3 standard errors of the oracle
    assert np.all(np.abs(p - pr) <= 3 * st)


def test_n64_plain_mc_mixed_matches_oracle(rb, orc):
    # computeIV_tensor's estimator on the N=64, M=2^16 parameters, split FP16 (k_fused, third stream)
    c = C1
    M = 2 ** 14
    Z = orc.normals_philox(C1_SEED, 0, M, 3, c["N"])
    pr, st = orc.price_stat(orc.payoff_mc_samples(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"],
                                                  c["N"], Z))
    for code in (rb.PREC_FP64, rb.PREC_SPLIT_FP16):
        p, _, s = rb.price_v2(c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], M, C1_SEED,
                              precision=code, method=rb.METHOD_MC)
        tol = 1e-12 if code == rb.PREC_FP64 else 1e-4
        assert np.max(np.abs(p - pr)) <= tol and np.max(np.abs(s - st)) <= tol


# ------------------------------------------------------------------- 256 parameter sets
def _requests(rb, ds, M_cap=None):
    from rbergomi_amd.serving import Request
    return [Request(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"],
                    min(d["M"], M_cap) if M_cap else d["M"], d["seed"]) for d in ds]


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_256_sets_batched_launch_within_budget_of_fp64(rb, bench, prec):
    from rbergomi_amd.serving import price_batch
    ds = bench.batched_requests()
    assert len(ds) == 256
    say("256 sets: %s batch vs FP64 one by one" % prec)
    reqs = price_batch(_requests(rb, ds), precision=getattr(rb, CODES[prec]))
    worst = 0.0
    for d, r in zip(ds, reqs):
        p64, _, s64 = rb.price_v2(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"])
        e = float(np.max(np.abs(r.price - p64)))
        worst = max(worst, e)
        assert e <= 1e-4, (d["N"], d["M"], e)
        assert np.max(np.abs(r.stat - s64)) <= 1e-4
    say("   worst |price - fp64| over 256 x 21 prices: %.3e" % worst)


SETS256_ORACLE_SETS = [0, 1, 2, 5, 6, 10, 15, 32]  # N = 64, 128, 256, 32, 512, 512, 128, 32


def test_256_sets_match_oracle(rb, orc, bench):
    # eight of the 256 sets (every N of the grid among them) directly against the oracle on the
    # same Philox counters, at M = 2^13; the other sets ride in the same batched launch
    from rbergomi_amd.serving import price_batch
    ds = bench.batched_requests()
    Ms = 2 ** 13
    for prec in ("fp16", "bf16", "fp32gemm"):
        reqs = price_batch(_requests(rb, ds, M_cap=Ms), precision=getattr(rb, CODES[prec]))
        for i in SETS256_ORACLE_SETS:
            d = ds[i]
            Z = orc.normals_philox(d["seed"], 0, Ms, 2, d["N"])
            pr, st = orc.price_stat(orc.payoff_rt_samples(d["xi"], d["H"], d["eta"], d["rho"], d["T"],
                                                          d["K"], d["N"], Z))
            e = float(np.max(np.abs(reqs[i].price - pr)))
            say("256 sets, set %d (N=%d) %s: |price - oracle| <= %.3e" % (i, d["N"], prec, e))
            assert e <= 1e-4
            assert np.max(np.abs(reqs[i].stat - st)) <= 1e-4
            if prec == "fp16":
                p64, _, s64 = rb.price_v2(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], Ms,
                                          d["seed"])
                assert np.max(np.abs(p64 - pr)) <= 1e-12 and np.max(np.abs(s64 - st)) <= 1e-12


# ------------------------------------------------------------------------- long paths
@pytest.mark.parametrize("N,M", [(1024, 4096), (2048, 2048), (4096, 1024)])
def test_long_paths_match_oracle(rb, orc, N, M):
    # split FP16: the Toeplitz GEMM kernel at N = 1024, the single-launch LDS-FFT kernel beyond
    K = list(np.linspace(0.7, 1.3, 31))
    T = [1.0] * 31
    for H in (0.02, 0.05):
        seed = [4, N, int(H * 100)]
        Z = orc.normals_philox(seed, 0, M, 2, N)
        pr, st = orc.price_stat(orc.payoff_rt_samples(0.03, [H], [2.3], [-0.9], T, K, N, Z))
        for code in (rb.PREC_SPLIT_FP16, rb.PREC_FP32_FFT, rb.PREC_FP64):
            p, _, s = rb.price_v2(0.03, [H], [2.3], [-0.9], T, K, N, M, seed, precision=code)
            e = float(np.max(np.abs(p - pr)))
            say("long paths N=%d H=%.2f prec=%d: |price - oracle| <= %.3e" % (N, H, code, e))
            tol = 1e-12 if code == rb.PREC_FP64 else 1e-4
            assert e <= tol and np.max(np.abs(s - st)) <= tol


@pytest.mark.parametrize("N", [512, 2048])
def test_fft_kernel_plain_mc_matches_oracle(rb, orc, N):
    # k_fused_fft with the plain-MC estimator (third Gaussian stream), RB_PREC_FP32_FFT at N = 512
    # and the mixed long-path kernel at N = 2048
    K = list(np.linspace(0.8, 1.2, 9))
    T = [0.5] * 4 + [1.0] * 5
    M = 2048
    seed = [17, N]
    Z = orc.normals_philox(seed, 0, M, 3, N)
    pr, st = orc.price_stat(orc.payoff_mc_samples(0.04, [0.1], [1.5], [-0.7], T, K, N, Z))
    code = rb.PREC_FP32_FFT if N <= 1024 else rb.PREC_SPLIT_FP16
    p, _, s = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, N, M, seed, precision=code, method=rb.METHOD_MC)
    assert np.max(np.abs(p - pr)) <= 1e-4 and np.max(np.abs(s - st)) <= 1e-4


# ---------------------------------------------------------------------- serving trace
def test_serving_trace_slice(rb, bench):
    """The first 384 requests of the 4096-request serving trace through the server with the
    benchmark's settings: each result is bit-identical to pricing the request alone, within
    1e-4 of the FP64 path, and the device timeline of the batches' kernels (timing mode) shows
    kernels of different batches running at the same time."""
    from rbergomi_amd.serving import Server
    ds = bench.serving_requests(384)
    assert [d["seed"] for d in ds] == [d["seed"] for d in bench.serving_requests(4096)[:384]]
    reqs = _requests(rb, ds)
    say("serving trace: 384 requests, closed loop")
    with Server(rb.PREC_SPLIT_FP16, max_inflight=bench.SERVE_INFLIGHT, max_batch_requests=bench.SERVE_MAX_BATCH,
                max_batch_path_steps=bench.SERVE_CAP) as srv:
        srv.set_timing(True)
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            srv.wait(r)
        rep = srv.report()
    say("   report:", rep)
    # several batches in flight at once, and their kernels' device intervals overlapping
    assert rep["completed"] == 384 and rep["peak_inflight"] >= 2 and rep["batches"] >= 8
    assert rep["kernel_launches"] == rep["batches"]
    assert rep["max_concurrent"] >= 2 and rep["device_overlap_ms"] > 0.0
    assert 0.0 < rep["device_busy_ms"] <= rep["kernel_ms"] * (1 + 1e-9)
    flops = sum(d["M"] * d["N"] * (d["N"] - 1) for d in ds)  # one H per request
    assert abs(rep["conv_flops"] - flops) <= 1e-9 * flops
    worst = 0.0
    for d, r in zip(ds, reqs):
        p, iv, s = rb.price_v2(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"],
                               precision=rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price) and np.array_equal(s, r.stat)
        assert np.array_equal(iv, r.iv, equal_nan=True)
    for d, r in list(zip(ds, reqs))[::4]:
        p64 = rb.price_v2(d["xi"], d["H"], d["eta"], d["rho"], d["T"], d["K"], d["N"], d["M"], d["seed"])[0]
        e = float(np.max(np.abs(r.price - p64)))
        worst = max(worst, e)
        assert e <= 1e-4
    say("   worst |price - fp64| over every 4th request: %.3e" % worst)
