"""GPU checks of unordered-row pricing (rb_price_rows, rb_payoff_rows_samples; the reference's
ComputeIVRTMTunstructured, rBergomi/rBergomiMT.cpp:837-935, with a per-row xi): per-path payoffs
against the oracle on caller Gaussians; seeded prices against the oracle on the device's Philox
Gaussians; the tensor form laid out as rows gives the structured results bit for bit (FP64 and
split precision, both estimators); split precision within 1e-4 of FP64; argument errors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


def random_rows(n, seed):
    rng = np.random.default_rng(seed)
    xi = rng.uniform(0.01, 0.16, n)
    H = rng.choice([0.05, 0.1, 0.3], n)          # repeated H values share a Volterra unit
    eta = rng.uniform(0.5, 3.0, n)
    rho = rng.uniform(-0.95, -0.1, n)
    T = rng.choice([0.1, 0.5, 1.0, 2.0], n)
    K = np.exp(rng.uniform(-0.4, 0.3, n))
    return xi, H, eta, rho, T, K


def tensor_rows(H, eta, rho, T, K):
    idx = [(h, t) for h in range(len(H)) for t in range(len(T))]
    return ([H[h] for h, _ in idx], [eta[h] for h, _ in idx], [rho[h] for h, _ in idx],
            [T[t] for _, t in idx], [K[t] for _, t in idx])


@pytest.mark.parametrize("N", [16, 100, 256])
def test_row_payoffs_match_oracle(rb, orc, N):
    xi, H, eta, rho, T, K = random_rows(13, N)
    Z = np.random.default_rng(N + 1).standard_normal((40, 2, N))
    print("rows payoffs N=%d" % N, flush=True)
    payoff, price, stat = rb.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z)
    ref = orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z)
    assert np.max(np.abs(payoff - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-10
    pr, st = orc.price_stat(ref)
    assert np.max(np.abs(price - pr)) < 1e-12


def test_seeded_row_prices_match_oracle(rb, orc):
    xi, H, eta, rho, T, K = random_rows(11, 3)
    N, M, seed = 48, 2500, [9, 4]
    price, iv, stat = rb.price_rows(xi, H, eta, rho, T, K, N, M, seed)
    Z = orc.normals_philox(seed, 0, M, 2, N)
    pr, st = orc.price_stat(orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z))
    assert np.max(np.abs(price - pr)) < 1e-12
    assert np.max(np.abs(stat - st)) < 1e-12
    ok = np.isfinite(iv)
    assert ok.sum() >= 8


@pytest.mark.parametrize("prec", [0, 2, 3])
@pytest.mark.parametrize("method", [0, 1])
def test_tensor_as_rows_equals_structured(rb, prec, method):
    H, eta, rho = [0.07, 0.1], [1.9, 1.5], [-0.9, -0.7]
    T, K = [0.25, 1.0, 1.0, 2.0], [0.9, 1.0, 1.2, 0.8]
    args = (256, 5000, [3, 1])
    a = rb.price_v2(0.04, H, eta, rho, T, K, *args, precision=prec, method=method)
    b = rb.price_rows(0.04, *tensor_rows(H, eta, rho, T, K), *args, precision=prec, method=method)
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_reduced_precision_rows_within_budget(rb, prec):
    xi, H, eta, rho, T, K = random_rows(40, 11)
    args = (xi, H, eta, rho, T, K, 128, 6000, [1, 2])
    p64 = rb.price_rows(*args)[0]
    p = rb.price_rows(*args, precision=prec)[0]
    assert np.max(np.abs(p - p64)) <= 1e-4


def test_row_argument_errors(rb):
    with pytest.raises(rb.RbError) as e:
        rb.price_rows([-0.1], [0.1], [1.0], [-0.5], [1.0], [1.0], 16, 100, [1])
    assert e.value.code == 1
    L = rb.load()
    dp = ctypes.POINTER(ctypes.c_double)
    one = np.ones(1)
    out = np.zeros(1)
    seed = np.ones(1, dtype=np.uint64)
    rc = L.rb_price_rows(*[one.ctypes.data_as(dp)] * 6, 0, 16, 100,
                         seed.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 1, None,
                         out.ctypes.data_as(dp), out.ctypes.data_as(dp), out.ctypes.data_as(dp))
    assert rc == 1


TURBO_PILOT_BASE = 1 << 62   # Philox path counters of the pilot paths


def turbo_case():
    K = np.array([0.8, 0.95, 1.0, 1.1, 1.3, 1.0])
    n = len(K)
    H = np.array([0.1, 0.1, 0.1, 0.1, 0.1, 0.3])
    eta = np.array([1.5, 1.5, 1.5, 1.5, 1.5, 0.9])
    rho = np.array([-0.7, -0.7, -0.7, -0.7, -0.7, -0.3])
    T = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.5])
    xi = np.array([0.04, 0.04, 0.04, 0.04, 0.04, 0.09])
    return xi, H, eta, rho, T, K


@pytest.mark.parametrize("mc,qmax", [(3000, 10000.0), (3, 50.0)])
def test_turbo_matches_oracle_on_same_gaussians(rb, orc, mc, qmax):
    """The device estimator against the oracle's restatement of ComputePriceRTVarRed on the same
    pilot and main Gaussians (the small pilot forces Q <= Iv exceptions)."""
    xi, H, eta, rho, T, K = turbo_case()
    N, M, seed = 32, 4000, [21, 5]
    print("turbo mc=%d" % mc, flush=True)
    got = rb.price_rows_turbo(xi, H, eta, rho, T, K, N, M, seed, m_control=mc, q_max=qmax)
    Zc = orc.normals_philox(seed, TURBO_PILOT_BASE, mc, 2, N)
    Z = orc.normals_philox(seed, 0, M, 2, N)
    ref = orc.turbo_rows(xi, H, eta, rho, T, K, N, Zc, Z, qmax=qmax)
    assert np.array_equal(got["exceptions"], ref["exceptions"])
    if mc == 3:
        assert np.all(got["exceptions"] > 0)
    assert np.max(np.abs(got["Q"] / ref["Q"] - 1)) < 1e-12
    assert np.max(np.abs(got["alpha"] - ref["alpha"])) < 1e-9 * np.max(np.abs(ref["alpha"]))
    assert np.max(np.abs(got["price"] - ref["price"])) < 1e-11
    assert np.max(np.abs(got["stat"] / ref["stat"] - 1)) < 1e-9


def test_turbo_agrees_with_plain_rt_and_reduces_variance(rb):
    xi, H, eta, rho, T, K = turbo_case()
    N, M = 64, 200000
    tb = rb.price_rows_turbo(xi, H, eta, rho, T, K, N, M, [1, 1])
    p, iv, s = rb.price_rows(xi, H, eta, rho, T, K, N, M, [2, 2])
    assert np.all(np.abs(tb["price"] - p) <= 4.5 * np.sqrt(tb["stat"] ** 2 + s ** 2))
    assert np.all(tb["stat"] < s)
    assert np.all(np.isfinite(tb["iv"]))


def test_turbo_argument_errors(rb):
    xi, H, eta, rho, T, K = turbo_case()
    with pytest.raises(rb.RbError) as e:
        rb.price_rows_turbo(xi, H, eta, rho, T, K, 32, 1000, [1], m_control=1)
    assert e.value.code == 1
    with pytest.raises(rb.RbError) as e:
        rb.price_rows_turbo(xi, H, eta, rho, T, K, 32, 1000, [1], q_max=-1.0)
    assert e.value.code == 1
