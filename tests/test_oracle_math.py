"""Pins of the CPU oracle's building blocks against mathematics written independently here
(quadrature, exact rationals, textbook identities, published known-answer vectors).

Reference functions covered: getGamma (rBergLib/rBergomiMT.cpp:143-149), the FFT
convolution of updateWtilde (:83-109), Phi / BS_call_price / BS_put_price / IV_call
(rBergLib/BlackScholes.cpp:16-92)."""
from fractions import Fraction

import numpy as np
import pytest
from scipy import integrate, stats


@pytest.mark.parametrize("H", [0.02, 0.07, 0.1, 0.3, 0.45])
def test_gamma_is_cell_integral_of_kernel(orc, H):
    # Gamma_k = int_k^{k+1} s^(H-1/2) ds (the hybrid-scheme weight b*_{k+1}^alpha),
    # checked by adaptive quadrature instead of the closed form.
    N = 64
    g = orc.gamma(H, N)
    assert g[0] == 0.0
    for k in [1, 2, 3, 10, 63]:
        val, _ = integrate.quad(lambda s: s ** (H - 0.5), k, k + 1, epsabs=0, epsrel=1e-13)
        assert abs(g[k] - val) <= 1e-12 * abs(val)


@pytest.mark.parametrize("H", [0.05, 0.25])
def test_gamma_telescoping_sum(orc, H):
    N = 2048
    g = orc.gamma(H, N)
    a1 = H + 0.5
    i = np.arange(1, N)
    want = ((i + 1.0) ** a1 - 1.0) / a1
    assert np.allclose(np.cumsum(g[1:]), want, rtol=1e-12, atol=0)


def test_gamma_half_is_flat(orc):
    g = orc.gamma(0.5, 100)
    assert g[0] == 0 and np.all(np.abs(g[1:] - 1.0) < 1e-15)


def test_convolution_reference_known_answer(orc):
    # rBergomi/text/tests.cpp:333-361 (commented-out test of the reference): z1 = x * y
    x = [1.0, 2.4, 3.9, 4.8]
    y = [9.0, 5.4, 2.9, 7.8]
    want = [9.0, 27.0, 50.96, 79.02]
    for fft in (False, True):
        z = orc.convolve(x, y, fft=fft)
        assert np.allclose(z, want, rtol=1e-13, atol=1e-13)


def test_convolution_exact_rationals(orc):
    rng = np.random.default_rng(5)
    n = 37
    xi = rng.integers(-50, 50, n)
    yi = rng.integers(-50, 50, n)
    x = [Fraction(int(v), 8) for v in xi]
    y = [Fraction(int(v), 16) for v in yi]
    exact = [sum(x[k] * y[i - k] for k in range(i + 1)) for i in range(n)]
    z = orc.convolve([float(v) for v in x], [float(v) for v in y])
    assert all(float(e) == zz for e, zz in zip(exact, z))  # dyadic inputs: exact in FP64
    zf = orc.convolve([float(v) for v in x], [float(v) for v in y], fft=True)
    assert np.allclose(zf, [float(e) for e in exact], rtol=0, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 5, 64, 255, 256, 1000])
def test_fft_equals_direct(orc, n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    assert np.allclose(orc.convolve(x, y, fft=True), orc.convolve(x, y), rtol=0, atol=1e-11 * n)


@pytest.mark.parametrize("H,N", [(0.07, 16), (0.1, 64), (0.4, 33)])
def test_wtilde_impulse_responses(orc, H, N):
    """W1 = e_j gives the hybrid-scheme weights; W1perp = e_j gives the first-cell term only.
    Weights written from the scheme (Bennedsen-Lunde-Pakkanen, kappa = 1): far cells
    int_k^{k+1} s^(H-1/2) ds, near cell correlated through rho_H = sqrt(2H)/(H+1/2)."""
    sc = np.sqrt(2 * H) * N ** (-H)
    rhoH = np.sqrt(2 * H) / (H + 0.5)
    a = rhoH / np.sqrt(2 * H)
    b = np.sqrt(1 - rhoH ** 2) / np.sqrt(2 * H)
    far = [0.0] + [integrate.quad(lambda s: s ** (H - 0.5), k, k + 1, epsrel=1e-13)[0] for k in range(1, N)]
    for j in [0, 1, N // 2, N - 1]:
        e = np.zeros(N); e[j] = 1.0
        for conv in (orc.CONV_DIRECT, orc.CONV_FFT):
            w = orc.wtilde(e, np.zeros(N), H, conv)
            want = np.array([sc * far[i - j] if i > j else (sc * a if i == j else 0.0) for i in range(N)])
            assert np.allclose(w, want, rtol=1e-11, atol=1e-13)
            wp = orc.wtilde(np.zeros(N), e, H, conv)
            wantp = np.zeros(N); wantp[j] = sc * b
            assert np.allclose(wp, wantp, rtol=1e-12, atol=1e-13)


def test_wtilde_variance_closed_form(orc):
    # Var(Wtilde_i) = N^(-2H) (1 + 2H sum_{k<=i} Gamma_k^2); near cell has variance N^(-2H).
    H, N = 0.1, 32
    cols = []
    for j in range(N):
        e = np.zeros(N); e[j] = 1.0
        cols.append(orc.wtilde(e, np.zeros(N), H))
        cols.append(orc.wtilde(np.zeros(N), e, H))
    A = np.array(cols)  # rows: basis inputs, cols: outputs
    var = (A ** 2).sum(axis=0)
    g = orc.gamma(H, N)
    want = N ** (-2 * H) * (1 + 2 * H * np.cumsum(g ** 2))
    assert np.allclose(var, want, rtol=1e-12)
    assert abs(var[0] - N ** (-2 * H)) < 1e-13


def test_update_v_reference_compensator(orc):
    # v_0 = xi; v_1 uses Wtilde_0 with a zero compensator ((1-1)*dt)^(2H) = 0; Wtilde_{N-1} unused
    N, xi, H, eta, T = 8, 0.04, 0.1, 1.5, 2.0
    Wt = np.linspace(-0.3, 0.4, N)
    v = orc.update_v(Wt, xi, H, eta, T)
    assert v[0] == xi
    assert abs(v[1] - xi * np.exp(eta * T ** H * Wt[0])) < 1e-15
    Wt2 = Wt.copy(); Wt2[-1] = 123.0
    assert np.array_equal(orc.update_v(Wt2, xi, H, eta, T), v)
    vx = orc.update_v(Wt, xi, H, eta, T, orc.COMP_EXACT)
    assert vx[0] == xi and vx[1] < v[1]


def test_phi_matches_scipy(orc):
    for x in [-38.0, -10.0, -3.3, -1.0, 0.0, 0.5, 2.0, 8.0]:
        assert abs(orc.phi(x) - stats.norm.cdf(x)) <= 1e-15 + 1e-13 * stats.norm.cdf(x)


def test_bs_identities(orc):
    for sigma in [0.05, 0.2, 0.8]:
        atm = orc.bs_call(1.0, 1.0, 1.0, sigma)
        assert abs(atm - (2 * stats.norm.cdf(sigma / 2) - 1)) < 1e-15
        for K in [0.7, 1.0, 1.3]:
            for tau in [0.1, 1.0, 2.0]:
                c, p = orc.bs_call(1.0, K, tau, sigma), orc.bs_put(1.0, K, tau, sigma)
                assert abs((c - p) - (1.0 - K)) < 1e-14


@pytest.mark.parametrize("K", [0.5, 0.8, 0.95, 1.0, 1.05, 1.2, 1.6])
@pytest.mark.parametrize("tau", [0.05, 0.5, 1.0, 2.0])
def test_iv_round_trip(orc, K, tau):
    for sigma in [0.05, 0.12, 0.3, 0.7, 1.5]:
        price = orc.bs_call(1.0, K, tau, sigma)
        d1 = (np.log(1.0 / K) + 0.5 * sigma * sigma * tau) / (sigma * np.sqrt(tau))
        vega = stats.norm.pdf(d1) * np.sqrt(tau)
        if vega < 1e-12:
            continue  # price carries no information about sigma in double precision
        # condition number of the inverse: a price rounding of eps*price moves sigma by eps*price/vega
        tol = 1e-10 * sigma + 8 * np.finfo(float).eps * max(price, 1.0) / vega
        iv = orc.iv_call(price, 1.0, K, tau)
        assert abs(iv - sigma) <= tol, (K, tau, sigma, iv, tol)


def test_iv_outside_range_is_nan(orc):
    assert np.isnan(orc.iv_call(1e-3 * 0, 1.0, 1.2, 1.0))
    assert np.isnan(orc.iv_call(0.1, 1.0, 0.8, 1.0))  # below intrinsic 0.2
    assert np.isnan(orc.iv_call(1.5, 1.0, 1.0, 1.0))  # above S0


def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        assert list(orc.philox(ctr, key)) == want


def test_normals_are_box_muller_of_philox(orc):
    seed = [7, 11, 13]
    key = orc.philox_key(seed)
    z = orc.normals_philox(seed, 5, 3, 2, 10)
    for pp in range(3):
        p = 5 + pp
        for s in range(2):
            for q in range(3):
                x = orc.philox([q, p & 0xFFFFFFFF, p >> 32, s], key).astype(np.float64)
                u = (x + 0.5) / 2.0 ** 32
                r0 = np.sqrt(-2 * np.log(u[0])); r1 = np.sqrt(-2 * np.log(u[2]))
                want = [r0 * np.cos(2 * np.pi * u[1]), r0 * np.sin(2 * np.pi * u[1]),
                        r1 * np.cos(2 * np.pi * u[3]), r1 * np.sin(2 * np.pi * u[3])]
                got = z[pp, s, 4 * q:4 * q + 4]
                assert np.allclose(got, want[:len(got)], rtol=1e-14, atol=1e-15)


def test_normals_moments(orc):
    z = orc.normals_philox([2024], 0, 4000, 2, 256).ravel()
    n = z.size
    assert abs(z.mean()) < 4.5 / np.sqrt(n)
    assert abs(z.var() - 1) < 4.5 * np.sqrt(2.0 / n)
    assert abs(stats.kurtosis(z)) < 4.5 * np.sqrt(24.0 / n)
    # independence across paths and streams: correlation of neighbours
    zz = orc.normals_philox([2024], 0, 4000, 2, 256)
    c = np.corrcoef(zz[:, 0, :].ravel(), zz[:, 1, :].ravel())[0, 1]
    assert abs(c) < 4.5 / np.sqrt(zz[:, 0, :].size)


def test_iv_reference_tolerance_pins(orc):
    """The reference-tolerance IV (Boost's bracket from 0.2 by the factor 2, midpoint of a bracket
    of relative width 2^-10, rBergLib/BlackScholes.cpp:36-92): within 2^-11 relative of the
    exact IV of the same price; a price of BS(0.2 * 2^k) sits on a bracket end and returns it
    exactly; a vol the 20 evaluations cannot bracket gives NaN (Boost throws there)."""
    rng = np.random.default_rng(5)
    for _ in range(400):
        K = float(np.exp(rng.uniform(-0.5, 0.4)))
        T = float(rng.uniform(0.05, 2.0))
        sig = float(np.exp(rng.uniform(np.log(0.02), np.log(2.5))))
        price = orc.bs_call(1.0, K, T, sig)
        otm = price if K >= 1.0 else orc.bs_put(1.0, K, T, sig)  # what either solver inverts
        if otm < 1e-8:  # time value lost to rounding: no solver recovers the vol from such a price
            continue
        exact = orc.iv_call(price, 1.0, K, T)
        ref = orc.iv_call_reference(price, 1.0, K, T)
        assert abs(exact / sig - 1) < 1e-9
        assert abs(ref - exact) <= 2.0 ** -11 * exact * (1 + 1e-9), (K, T, sig, ref, exact)
    for k in (-3, -1, 0, 1, 3):
        s = 0.2 * 2.0 ** k
        for K in (1.0, 1.3):  # the call branch: f(0.2 * 2^k) is exactly zero there
            assert orc.iv_call_reference(orc.bs_call(1.0, K, 1.0, s), 1.0, K, 1.0) == s
    # sigma = 1e-7 lies 21 halvings below the guess: no bracket within 20 evaluations
    assert np.isnan(orc.iv_call_reference(orc.bs_call(1.0, 1.0, 1.0, 1e-7), 1.0, 1.0, 1.0))
    # but 1e-4 (11 halvings) is bracketed
    assert abs(orc.iv_call_reference(orc.bs_call(1.0, 1.0, 1.0, 1e-4), 1.0, 1.0, 1.0) / 1e-4 - 1) < 2.0 ** -10
    # outside the no-arbitrage range: NaN, both branches
    assert np.isnan(orc.iv_call_reference(1.2, 1.0, 1.1, 1.0))
    assert np.isnan(orc.iv_call_reference(0.1, 1.0, 0.8, 1.0))
