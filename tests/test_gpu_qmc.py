"""GPU checks of the quasi-Monte Carlo source (rb_price_sobol; ComputeIVRTMT_sobol,
rBergomi/rBergomiMT.cpp:610-716): the device inverse normal CDF against scipy.special.ndtri; prices
against the oracle on the same Sobol' points (host points -> ndtri -> breakZ -> oracle payoffs);
the QMC estimate against a large Philox reference; split precision within 1e-4 of FP64."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


def test_device_inverse_normal(rb):
    from scipy.special import ndtri
    u = np.concatenate([np.arange(1, 2 ** 16) / 2.0 ** 16, [2.0 ** -32, 1 - 2.0 ** -32, 2.0 ** -33,
                                                             1e-12, 0.5, 0.5 + 2.0 ** -32]])
    z = np.zeros_like(u)
    dp = ctypes.POINTER(ctypes.c_double)
    assert rb.load().rb_probe_norm_inv(u.ctypes.data_as(dp), len(u), z.ctypes.data_as(dp)) == 0
    ref = ndtri(u)
    assert np.max(np.abs(z - ref) / np.maximum(np.abs(ref), 1.0)) < 3e-15


def sobol_gaussians(rb, M, N, ns):
    from scipy.special import ndtri
    u = rb.sobol_points(ns * N, 1, M)           # path m uses point m + 1
    z = ndtri(u)
    # dimension ns*j + s -> stream s, step j (breakZ for two streams)
    return z.reshape(M, N, ns).transpose(0, 2, 1).copy()


@pytest.mark.parametrize("method", [0, 1])
def test_sobol_prices_match_oracle(rb, orc, method):
    H, eta, rho = [0.07, 0.3], [1.9, 0.9], [-0.9, -0.4]
    T, K = [0.5, 1.0], [0.95, 1.1]
    N, M = 32, 4000
    p, iv, s = rb.price_sobol(0.04, H, eta, rho, T, K, N, M, method=method)
    Z = sobol_gaussians(rb, M, N, 3 if method else 2)
    pay = (orc.payoff_mc_samples if method else orc.payoff_rt_samples)(0.04, H, eta, rho, T, K, N, Z)
    pr, st = orc.price_stat(pay)
    assert np.max(np.abs(p - pr)) < 1e-12
    assert np.max(np.abs(s - st)) < 1e-12


def shifted_sobol_gaussians(rb, orc, M, N, ns, seed):
    """points 0..M-1 with the seeded digital shift: shift word of dimension d = first output of
    Philox block (counter {0, d, 0, 7}, key of the seed), u = ((x ^ shift) + 1/2) / 2^32"""
    from scipy.special import ndtri
    dims = ns * N
    x = np.round(rb.sobol_points(dims, 0, M) * 2.0 ** 32).astype(np.uint64)
    key = orc.philox_key(seed)
    shift = np.array([orc.philox([0, d, 0, 7], key)[0] for d in range(dims)], dtype=np.uint64)
    u = ((x ^ shift[None, :]).astype(np.float64) + 0.5) / 2.0 ** 32
    return ndtri(u).reshape(M, N, ns).transpose(0, 2, 1).copy()


def test_shifted_sobol_prices_match_oracle(rb, orc):
    H, eta, rho = [0.07, 0.3], [1.9, 0.9], [-0.9, -0.4]
    T, K = [0.5, 1.0], [0.95, 1.1]
    N, M, seed = 24, 3000, [4, 2]
    p, iv, s = rb.price_sobol(0.04, H, eta, rho, T, K, N, M, seed=seed)
    Z = shifted_sobol_gaussians(rb, orc, M, N, 2, seed)
    pr, st = orc.price_stat(orc.payoff_rt_samples(0.04, H, eta, rho, T, K, N, Z))
    assert np.max(np.abs(p - pr)) < 1e-12


def test_sobol_accuracy_against_large_philox_reference(rb):
    """The unscrambled points 1..M (the reference's construction) leave out the origin of every
    net and add one far-tail point per dimension: a bias of the order of the MC error at this M.
    The digitally shifted estimate is unbiased: independent shifts average to the reference."""
    args = (0.04, [0.1], [1.5], [-0.7], [0.5, 1.0, 1.0], [1.0, 0.9, 1.1], 32)
    ref, _, sref = rb.price_v2(*args, 1 << 22, [7, 7])
    q, _, sq = rb.price_sobol(*args, 1 << 16)
    print("unscrambled err", np.abs(q - ref), "mc se at 2^16", sq, flush=True)
    assert np.all(np.abs(q - ref) <= 4.5 * sref + 2.0 * sq)
    est = np.array([rb.price_sobol(*args, 1 << 14, seed=[100 + k])[0] for k in range(8)])
    mean, se = est.mean(axis=0), est.std(axis=0, ddof=1) / np.sqrt(len(est))
    mc_se = rb.price_v2(*args, 8 << 14, [9, 9])[2]
    print("rqmc err", np.abs(mean - ref), "rqmc se", se, "mc se same paths", mc_se, flush=True)
    assert np.all(np.abs(mean - ref) <= 4.5 * np.sqrt(se ** 2 + sref ** 2))


def test_sobol_split_precision_within_budget(rb):
    args = (0.04, [0.1], [1.5], [-0.7], [0.5, 1.0], [1.0, 1.1], 256, 20000)
    p64 = rb.price_sobol(*args)[0]
    p16 = rb.price_sobol(*args, precision=rb.PREC_SPLIT_FP16)[0]
    assert np.max(np.abs(p16 - p64)) <= 1e-4


# ---- Brownian-bridge mapping of the Sobol' points (rb_price_sobol_bridge)

@pytest.mark.parametrize("m", [0, 1, 4, 9, 13])
def test_device_bridge_matches_oracle(rb, orc, m):
    Z = np.random.default_rng(m).standard_normal((7, 1 << m))
    dW = rb.bridge_increments(Z)
    assert np.max(np.abs(dW - orc.bridge(Z))) < 1e-13


@pytest.mark.parametrize("method", [0, 1])
def test_bridged_sobol_prices_match_oracle(rb, orc, method):
    # host points -> ndtri -> per stream (breakZ order) -> bridge -> oracle payoffs
    H, eta, rho = [0.07, 0.3], [1.9, 0.9], [-0.9, -0.4]
    T, K = [0.5, 1.0], [0.95, 1.1]
    N, M = 32, 3000
    ns = 3 if method else 2
    p, iv, s = rb.price_sobol(0.04, H, eta, rho, T, K, N, M, method=method, bridge=True)
    Z = sobol_gaussians(rb, M, N, ns)
    Zb = orc.bridge(Z.reshape(M * ns, N)).reshape(M, ns, N)
    pay = (orc.payoff_mc_samples if method else orc.payoff_rt_samples)(0.04, H, eta, rho, T, K, N, Zb)
    pr, st = orc.price_stat(pay)
    assert np.max(np.abs(p - pr)) < 1e-12
    assert np.max(np.abs(s - st)) < 1e-12


def test_bridged_rqmc_is_unbiased(rb):
    args = (0.04, [0.1], [1.5], [-0.7], [0.5, 1.0, 1.0], [1.0, 0.9, 1.1], 64)
    ref, _, sref = rb.price_v2(*args, 1 << 22, [7, 8])
    est = np.array([rb.price_sobol(*args, 1 << 14, seed=[s, 1], bridge=True)[0] for s in range(8)])
    mean, se = est.mean(0), est.std(0, ddof=1) / np.sqrt(len(est))
    assert np.all(np.abs(mean - ref) <= 4.5 * np.sqrt(se ** 2 + sref ** 2))


def test_bridge_needs_power_of_two_steps(rb):
    with pytest.raises(rb.RbError) as e:
        rb.price_sobol(0.04, [0.1], [1.5], [-0.7], [1.0], [1.0], 48, 1000, bridge=True)
    assert e.value.code == 1
