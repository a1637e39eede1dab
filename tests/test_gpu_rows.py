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
