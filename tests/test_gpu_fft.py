"""GPU checks of the FP32 LDS-FFT Volterra convolution (long paths): per-path integrals against
the CPU oracle, and prices against the FP64 path on the same Philox counters."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


@pytest.mark.parametrize("N", [7, 100, 1100, 2048, 3000, 4096, 8192])
def test_fft_integrals_match_oracle(rb, orc, N):
    from rbergomi_amd._lib import RbConfig
    L = rb.load()
    H, eta, rho = [0.02, 0.1], [2.3, 1.5], [-0.9, -0.7]
    T, K = [1.0, 0.25], [1.0, 0.9]
    xi = 0.03
    M = 6 if N >= 4096 else 11
    Z = np.random.default_rng(N).standard_normal((M, 2, N))
    a = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    arrs = [a(x) for x in (H, eta, rho, T, K)]
    out = np.zeros((M, 4, 2))
    cfg = RbConfig(rb.PREC_FP32, rb.COMP_REF, rb.METHOD_RT, 0)
    rc = L.rb_probe_integrals(xi, arrs[0].ctypes.data_as(dp), arrs[1].ctypes.data_as(dp), arrs[2].ctypes.data_as(dp),
                              2, arrs[3].ctypes.data_as(dp), arrs[4].ctypes.data_as(dp), 2, N, M,
                              Z.ctypes.data_as(dp), ctypes.byref(cfg), out.ctypes.data_as(dp))
    assert rc == 0, L.rb_last_error()
    ref = orc.integrals_rt_samples(xi, H, eta, rho, T, K, N, Z)
    # FP32 transform of length L ~ 2N: |dY| ~ 1e-6 |Y|; Iv relative, Is against its scale
    iv_err = np.max(np.abs(out[..., 0] - ref[..., 0]) / ref[..., 0])
    is_err = np.max(np.abs(out[..., 1] - ref[..., 1]) / np.sqrt(ref[..., 0]))
    print("N=%d Iv rel %.2e  Is/sqrt(Iv) %.2e" % (N, iv_err, is_err))
    assert iv_err < 2e-6 and is_err < 2e-6


@pytest.mark.parametrize("N", [1500, 4096])
def test_fft_prices_match_fp64(rb, N):
    K = np.linspace(0.7, 1.3, 31)
    T = [1.0] * 31
    p64, _, s64 = rb.price_v2(0.03, [0.05], [2.3], [-0.9], T, K, N, 3000, [8], precision=rb.PREC_FP64)
    p32, _, s32 = rb.price_v2(0.03, [0.05], [2.3], [-0.9], T, K, N, 3000, [8], precision=rb.PREC_FP32)
    err = np.max(np.abs(p32 - p64))
    print("N=%d max |price fp32-fft - fp64| = %.2e" % (N, err))
    assert err <= 1e-4


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("N,M", [(1100, 1999), (2048, 1500), (4096, 700), (8192, 301)])
def test_fused_long_path_prices_match_fp64(rb, N, M, method):
    # the mixed-precision precisions beyond N = 1024 (device Gaussians; Romano-Touzi and plain
    # Monte Carlo) run the single-launch FFT kernel: same Philox counters as the FP64 path, FP32
    # transform and integrals, FP64 sums; two units (H), several maturities and correlations, odd M
    H, eta, rho = [0.05, 0.2], [2.3, 1.2], [-0.9, -0.4]
    T = [0.25, 1.0, 1.0, 2.0, 0.5]
    K = [0.9, 0.8, 1.2, 1.0, 1.05]
    print("fused long path N=%d M=%d method=%d" % (N, M, method), flush=True)
    p64, iv64, s64 = rb.price_v2(0.03, H, eta, rho, T, K, N, M, [6, N], precision=rb.PREC_FP64, method=method)
    for code in (rb.PREC_SPLIT_FP16, rb.PREC_FP32_GEMM):
        pm, ivm, sm = rb.price_v2(0.03, H, eta, rho, T, K, N, M, [6, N], precision=code, method=method)
        err = np.max(np.abs(pm - p64))
        print("   precision %d: max |price - fp64| = %.2e, |stat - fp64| = %.2e" % (code, err, np.max(np.abs(sm - s64))),
              flush=True)
        assert err <= 1e-4 and np.max(np.abs(sm - s64)) <= 1e-4
        assert np.all(np.isfinite(ivm) | ~np.isfinite(iv64))


def test_fused_long_path_reruns_and_plan_timing(rb):
    T, K = [1.0] * 5, list(np.linspace(0.8, 1.2, 5))
    a = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, 2048, 3001, [2], precision=rb.PREC_SPLIT_FP16)
    b = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, 2048, 3001, [2], precision=rb.PREC_SPLIT_FP16)
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)
    pl = rb.Plan(0.04, [0.1], [1.5], [-0.7], T, K, 2048, 3001, [2], precision=rb.PREC_SPLIT_FP16, profile=True)
    pl.run()
    times = pl.kernel_times()
    p, _, _ = pl.results()
    pl.close()
    assert "k_fused_fft" in times and np.array_equal(p, a[0])


@pytest.mark.parametrize("N,M", [(1025, 1), (1025, 2), (2049, 3), (4097, 5)])
def test_fused_long_path_edges(rb, N, M):
    # the first long N of each transform length, a single path (the pair's second path absent)
    T, K = [0.5, 1.0], [1.0, 1.1]
    for method in (0, 1):
        p64, _, s64 = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, N, M, [3, M], method=method)
        pm, _, sm = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, N, M, [3, M], method=method,
                                precision=rb.PREC_SPLIT_FP16)
        assert np.all(np.isfinite(pm))
        assert np.max(np.abs(pm - p64)) <= 1e-4


def test_fused_long_path_rejects_what_lds_cannot_hold(rb):
    # per-option FP64 sums live in LDS beside the L = 16384 transform at N = 8192
    K = list(np.linspace(0.5, 1.5, 4000))
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.1], [1.5], [-0.7], [1.0] * len(K), K, 8192, 10, [1], precision=rb.PREC_SPLIT_FP16)
    assert e.value.code == 5


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("N,M", [(7, 100), (32, 999), (100, 1501), (512, 2001), (513, 777), (1000, 600),
                                 (1024, 301), (2048, 500)])
def test_fft_precision_prices_match_fp64(rb, N, M, method):
    # PREC_FP32_FFT runs the single-launch FFT kernel at every N (transforms of 2^10 and 2^11 below
    # N = 1024, one quad per thread): same Philox counters as the FP64 path, FP32 transform and
    # integrals, FP64 sums; two units, several maturities and correlations, odd M
    H, eta, rho = [0.07, 0.3], [1.9, 1.1], [-0.9, -0.3]
    T = [0.1, 0.5, 1.0, 2.0]
    K = [0.95, 0.85, 1.2, 1.0]
    print("fft precision N=%d M=%d method=%d" % (N, M, method), flush=True)
    p64, iv64, s64 = rb.price_v2(0.04, H, eta, rho, T, K, N, M, [8, N], precision=rb.PREC_FP64, method=method)
    pm, ivm, sm = rb.price_v2(0.04, H, eta, rho, T, K, N, M, [8, N], precision=rb.PREC_FP32_FFT, method=method)
    err = np.max(np.abs(pm - p64))
    print("   max |price - fp64| = %.2e, |stat - fp64| = %.2e" % (err, np.max(np.abs(sm - s64))), flush=True)
    assert err <= 1e-4 and np.max(np.abs(sm - s64)) <= 1e-4
    assert np.all(np.isfinite(ivm) | ~np.isfinite(iv64))


def test_fft_precision_routing_and_batches(rb):
    # the plan runs the single-launch FFT kernel at N = 512; caller Gaussians take the staged FFT
    # pipeline; batches and the server run the Toeplitz-GEMM kernel only and reject the precision
    T, K = [1.0] * 3, [0.9, 1.0, 1.1]
    pl = rb.Plan(0.04, [0.1], [1.5], [-0.7], T, K, 512, 2000, [4], precision=rb.PREC_FP32_FFT, profile=True)
    pl.run()
    times = pl.kernel_times()
    p, _, _ = pl.results()
    pl.close()
    assert "k_fused_fft" in times and "k_fused" not in times
    again, _, _ = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, 512, 2000, [4], precision=rb.PREC_FP32_FFT)
    assert np.array_equal(p, again)
    from rbergomi_amd.serving import Request, price_batch
    req = Request(0.04, [0.1], [1.5], [-0.7], T, K, 512, 1000, [1])
    with pytest.raises(rb.RbError) as e:
        price_batch([req], precision=rb.PREC_FP32_FFT)
    assert e.value.code == 1
