"""GPU checks of the fused mixed-precision kernel (FP16 / BF16 hi+lo split and FP32 operands).

- the MFMA operand pairing and output-row layout it relies on (exact small-integer product);
- prices against the FP64 path on identical Philox counters: <= 1e-4 absolute (the stated
  mixed-precision budget); a layout or indexing mistake shows up as O(1e-2) errors;
- bit-identical reruns (deterministic reduction order).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mfma_layout(rb, mode):
    rng = np.random.default_rng(3 + mode)
    A = rng.integers(-2, 3, (32, 16)).astype(np.float32)
    B = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    out = np.zeros(1024, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    assert rb.load().rb_probe_mfma_layout(A.ctypes.data_as(fp), B.ctypes.data_as(fp), mode,
                                          out.ctypes.data_as(fp)) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    for lane in range(64):
        for r in range(16):
            row = 8 * (r // 4) + 4 * (lane // 32) + (r % 4)
            assert out[lane * 16 + r] == D[row, lane % 32], (lane, r)


def test_mfma_f64_layout(rb):
    # v_mfma_f64_16x16x4_f64 as the FP64 convolution uses it: lane l supplies A[l%16][l/16] and
    # B[l/16][l%16] and holds D[4 r + l/16][l%16]
    rng = np.random.default_rng(7)
    A = np.zeros(512, dtype=np.float32)
    B = np.zeros(512, dtype=np.float32)
    A[:64] = rng.integers(-3, 4, 64)
    B[:64] = rng.integers(-3, 4, 64)
    out = np.zeros(1024, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    assert rb.load().rb_probe_mfma_layout(A.ctypes.data_as(fp), B.ctypes.data_as(fp), 3,
                                          out.ctypes.data_as(fp)) == 0
    D = A[:64].reshape(16, 4).astype(np.float64) @ B[:64].reshape(4, 16).astype(np.float64)
    for lane in range(64):
        for r in range(4):
            assert out[lane * 4 + r] == D[4 * r + lane // 16, lane % 16], (lane, r)


def opts_config2_like():
    mats = [0.1, 0.25, 0.5, 1.0, 2.0]
    ks = np.linspace(-0.4, 0.2, 41)
    return np.repeat(mats, 41), np.exp(np.tile(ks, 5))


CASES = [
    # (N, M, H, eta, rho, T, K)
    (512, 4096, [0.1], [1.5], [-0.7], None, None),
    (256, 3000, [0.07, 0.1, 0.07], [1.9, 1.5, 2.5], [-0.9, -0.7, -0.5], [0.1, 0.5, 0.5, 2.0, 1.0],
     [0.9, 0.8, 1.25, 1.0, 1.6]),
    (64, 2000, [0.07], [1.9], [-0.9], [1.0] * 21, list(np.linspace(0.8, 1.2, 21))),
    (32, 1500, [0.3], [0.8], [-0.3], [0.05, 2.0], [1.0, 0.7]),
    (96, 999, [0.02, 0.45], [4.0, 0.5], [-0.99, -0.1], [0.2, 1.0], [0.9, 1.1]),
    (200, 777, [0.07], [2.2], [-0.9], [0.5], [0.8]),
    (1000, 600, [0.05], [2.3], [-0.9], [1.0, 1.0], [0.7, 1.3]),
    (1024, 512, [0.02], [2.3], [-0.9], [1.0], [1.0]),
]


CODES = {"fp16": "PREC_SPLIT_FP16", "bf16": "PREC_SPLIT_BF16", "fp32": "PREC_FP32_GEMM"}


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_matches_fp64_on_same_counters(rb, case, prec):
    N, M, H, eta, rho, T, K = CASES[case]
    if T is None:
        T, K = opts_config2_like()
    xi = 0.04
    seed = [11, case]
    say("fused %s case %d N=%d M=%d" % (prec, case, N, M))
    p64, iv64, s64 = rb.price_v2(xi, H, eta, rho, T, K, N, M, seed, precision=rb.PREC_FP64)
    code = getattr(rb, CODES[prec])
    pm, ivm, sm = rb.price_v2(xi, H, eta, rho, T, K, N, M, seed, precision=code)
    err = np.max(np.abs(pm - p64))
    say("   max |price - fp64| = %.3e   max |stat - fp64| = %.3e" % (err, np.max(np.abs(sm - s64))))
    assert err <= 1e-4
    assert np.max(np.abs(sm - s64)) <= 1e-4


MC_CASES = [0, 1, 2, 3, 4, 6, 7]


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32"])
@pytest.mark.parametrize("case", MC_CASES)
def test_fused_plain_mc_matches_fp64_on_same_counters(rb, case, prec):
    # computeIV_tensor's estimator: ln S_T from a third Gaussian stream, payoff (S_T - K)^+
    N, M, H, eta, rho, T, K = CASES[case]
    if T is None:
        T, K = opts_config2_like()
    xi = 0.04
    seed = [13, case]
    say("fused MC %s case %d N=%d M=%d" % (prec, case, N, M))
    p64, iv64, s64 = rb.price_v2(xi, H, eta, rho, T, K, N, M, seed, precision=rb.PREC_FP64,
                                 method=rb.METHOD_MC)
    code = getattr(rb, CODES[prec])
    pm, ivm, sm = rb.price_v2(xi, H, eta, rho, T, K, N, M, seed, precision=code, method=rb.METHOD_MC)
    err = np.max(np.abs(pm - p64))
    say("   max |price - fp64| = %.3e   max |stat - fp64| = %.3e" % (err, np.max(np.abs(sm - s64))))
    assert np.all(np.isfinite(pm))
    assert err <= 1e-4
    assert np.max(np.abs(sm - s64)) <= 1e-4


@pytest.mark.parametrize("nt", [7, 8, 10, 13])
def test_fused_plain_mc_many_maturities(rb, nt):
    # 6 or 8 groups per pass over the accumulators (fewest passes), several passes for nt > 8
    T = list(np.repeat(np.linspace(0.05, 2.0, nt), 3))
    K = list(np.tile([0.8, 1.0, 1.25], nt))
    say("fused MC %d maturities" % nt)
    args = (0.04, [0.1, 0.3], [1.5, 0.9], [-0.7, -0.2], T, K, 128, 3000, [5, nt])
    p64, _, s64 = rb.price_v2(*args, precision=rb.PREC_FP64, method=rb.METHOD_MC)
    for code in (rb.PREC_SPLIT_FP16, rb.PREC_SPLIT_BF16, rb.PREC_FP32_GEMM):
        pm, _, sm = rb.price_v2(*args, precision=code, method=rb.METHOD_MC)
        assert np.max(np.abs(pm - p64)) <= 1e-4
        assert np.max(np.abs(sm - s64)) <= 1e-4


def test_fused_reruns_are_bit_identical(rb):
    T, K = opts_config2_like()
    a = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, 512, 5000, [1], precision=rb.PREC_SPLIT_FP16)
    b = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, 512, 5000, [1], precision=rb.PREC_SPLIT_FP16)
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


def test_split_precision_routes_outside_the_fused_kernel(rb):
    # N > 1024 runs the staged pipeline with the FP32 FFT convolution (plain Monte Carlo at
    # N = 2048 too); N = 64 plain Monte Carlo is the fused kernel
    T, K = [0.5, 1.0], [1.0, 1.1]
    for N, method in ((2048, rb.METHOD_RT), (2048, rb.METHOD_MC), (64, rb.METHOD_MC)):
        p64, _, s64 = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, N, 2000, [1], method=method)
        for code in (rb.PREC_SPLIT_FP16, rb.PREC_FP32_GEMM):
            pm, _, sm = rb.price_v2(0.04, [0.1], [1.5], [-0.7], T, K, N, 2000, [1], method=method,
                                    precision=code)
            assert np.max(np.abs(pm - p64)) <= 1e-4
    for code in (rb.PREC_SPLIT_FP16, rb.PREC_FP32_GEMM):
        with pytest.raises(rb.RbError) as e:
            rb.price_v2(0.04, [0.1], [1.5], [-0.7], [1.0], [1.0], 9000, 100, [1], precision=code)
        assert e.value.code == 5


@pytest.mark.parametrize("form", [0, 1])
def test_mixed_payoff_and_mills_ratio(rb, form):
    """The split-precision payoff (FP32 with one shared Gaussian factor and a fitted Mills
    ratio, FP64 normal CDFs below d1 = -8: form 0 through erfcx as the GEMM kernel evaluates it,
    form 1 through erfc as the long-path kernel does) against the FP64 Black-Scholes formula,
    and the Mills-ratio fit against scipy.special.erfcx."""
    from scipy.special import erfcx
    from scipy.stats import norm
    rng = np.random.default_rng(7)
    n = 20000
    lnS = rng.normal(-0.02, 0.3, n)
    sig = np.exp(rng.uniform(np.log(0.005), np.log(1.5), n))
    K = np.exp(rng.uniform(-0.6, 0.4, n))
    z = np.concatenate([np.linspace(0, 12, n // 2), np.exp(rng.uniform(-8, 8, n - n // 2))]).astype(np.float32)
    out = np.zeros(n)
    mills = np.zeros(n, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    assert rb.load().rb_probe_payoff(lnS.ctypes.data_as(dp), sig.ctypes.data_as(dp), K.ctypes.data_as(dp), n,
                                     out.ctypes.data_as(dp), z.ctypes.data_as(fp), mills.ctypes.data_as(fp),
                                     form) == 0
    zz = z.astype(np.float64)
    ref_m = erfcx(zz / np.sqrt(2)) * np.sqrt(np.pi / 2)
    assert np.max(np.abs(mills / ref_m - 1)) < 5e-7
    S = np.exp(lnS)
    d1 = (lnS - np.log(K)) / sig + 0.5 * sig
    ref = S * norm.cdf(d1) - K * norm.cdf(d1 - sig)
    err = np.abs(out - ref)
    # FP32 inputs: |d(ln S)| ~ 6e-8 |ln S| propagates through delta ~ S Phi(d1)
    tol = 2e-6 * np.maximum(S, K) * np.maximum(1e-3, norm.pdf(d1) / sig) + 1e-7 * np.maximum(S, K)
    assert np.all(err <= tol + 1e-300), np.max(err / tol)
    deep = d1 < -8
    assert deep.any() and np.all(np.abs(out[deep] - ref[deep]) <= 1e-8 * ref[deep] + 1e-300)
    # and against 40-digit arithmetic on some deep points (the FP64 difference above cancels to
    # ~3e-9 relative; the erfcx form does not)
    import mpmath as mp
    mp.mp.dps = 40
    for i in np.flatnonzero(deep)[:40]:
        e1 = (mp.mpf(lnS[i]) - mp.log(mp.mpf(K[i]))) / mp.mpf(sig[i]) + mp.mpf(sig[i]) / 2
        exact = mp.e ** mp.mpf(lnS[i]) * mp.ncdf(e1) - mp.mpf(K[i]) * mp.ncdf(e1 - mp.mpf(sig[i]))
        assert abs(out[i] - float(exact)) <= (1e-10 if form == 0 else 1e-8) * float(exact)


@pytest.mark.parametrize("prec,N", [("PREC_SPLIT_FP16", 45), ("PREC_SPLIT_BF16", 45), ("PREC_FP32_GEMM", 45),
                                    ("PREC_SPLIT_FP16", 64), ("PREC_SPLIT_FP16", 1100), ("PREC_FP32_FFT", 45)])
def test_flat_variance_limit_mixed(rb, prec, N):
    """eta = rho = 0: v = xi on every step, so every path's conditioned payoff is
    C_BS(1, K, T, sqrt(xi T)) and the price equals it up to FP32 rounding.  With eta = 0 the
    group coefficient alpha is the smallest positive float, and the steps past N of a padded row
    block (N = 45: rows 45..64; the long-path kernel's steps past N = 1100; the FFT kernel at
    N = 45) are zeroed only through P = +inf: one row they failed to zero would add xi dt to the
    variance and move the price by ~5e-4 at N = 45."""
    from scipy import stats
    T, K, xi = [0.5, 2.0, 1.0, 0.1], [1.0, 1.2, 0.8, 1.05], 0.05
    say("flat variance limit %s N=%d" % (prec, N))
    price, iv, stat = rb.price_v2(xi, [0.1], [0.0], [0.0], T, K, N, 3000, [1], precision=getattr(rb, prec))
    for j in range(len(T)):
        s = np.sqrt(xi * T[j])
        d1 = (np.log(1 / K[j]) + 0.5 * s * s) / s
        want = stats.norm.cdf(d1) - K[j] * stats.norm.cdf(d1 - s)
        assert abs(price[j] - want) < 2e-6, (j, price[j], want)
        assert abs(iv[j] - np.sqrt(xi)) < 1e-4


def test_issue_rate_probe(rb):
    """rb_probe_issue: positive per-instruction issue costs, a transcendental dearer than an FP32
    FMA, and an FP32 FMA within a factor of two of the 64-lane / 32-lane-per-clock floor at the
    device's peak engine clock (2.4 GHz: 0.83 ns)."""
    say("issue-rate probe")
    ns = (ctypes.c_double * 4)()
    simds = ctypes.c_int()
    assert rb.load().rb_probe_issue(ns, ctypes.byref(simds)) == 0
    assert simds.value == 1024, simds.value   # MI355X: 256 CUs
    fma, exp2, mad, f64 = list(ns)
    assert all(v > 0 for v in ns), list(ns)
    assert exp2 > 1.5 * fma and mad > fma and f64 >= 0.9 * fma, list(ns)
    assert 0.5 < fma < 2.0, fma


@pytest.mark.parametrize("method", ["rt", "mc"])
def test_fused_large_request_tail_items(rb, method):
    """N = 32, M = 2^22: 32768 tiles of 128 paths in 4096 items of eight, so the request's last
    eighth of tiles is cut into items of one tile, handed out last.  Every path must be priced
    once: the FP16 price matches the FP64 path on the same counters (a dropped or repeated item
    would move it by about an eighth of the price or by its statistical error)."""
    say("large request, tail items, %s" % method)
    T, K = [0.1, 0.1, 0.1], [0.95, 1.0, 1.05]
    meth = rb.METHOD_RT if method == "rt" else rb.METHOD_MC
    args = (0.04, [0.1], [1.5], [-0.7], T, K, 32, 2 ** 22, [5, 1])
    p64, _, s64 = rb.price_v2(*args, precision=rb.PREC_FP64, method=meth)
    p16, _, s16 = rb.price_v2(*args, precision=rb.PREC_SPLIT_FP16, method=meth)
    assert np.max(np.abs(p16 - p64)) <= 1e-6, (p16, p64)
    assert np.max(np.abs(s16 - s64)) <= 1e-6
