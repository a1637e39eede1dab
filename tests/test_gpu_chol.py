"""GPU checks of the exact (Cholesky) scheme, RB_SCHEME_CHOLESKY (RfBm, PremiaRBergomi/RfBm.cpp):
per-path payoffs against the oracle's RfBm restatement on caller Gaussians (RT and plain MC),
seeded prices against the oracle on the device's Philox Gaussians, and the reference's recorded
prices (rBergomiR/test_package_deep.R:31-32: xi = 0.04, H = 0.06, eta = 2.5, rho = -0.8, T = 0.8,
K = 1, 1000 steps, plain MC with the exact compensator, 1e5 paths): Cholesky 0.05784862 +- 0.00011394
and hybrid 0.05761252 +- 0.0001136558."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


OPTS = dict(H=[0.07, 0.3], eta=[1.9, 0.9], rho=[-0.9, -0.4], T=[0.5, 1.0, 2.0], K=[0.9, 1.0, 1.2])


@pytest.mark.parametrize("N", [16, 100, 300])
@pytest.mark.parametrize("method", [0, 1])
def test_cholesky_payoffs_match_oracle(rb, orc, N, method):
    o = OPTS
    ns = 3 if method == 1 else 2
    Z = np.random.default_rng(N + method).standard_normal((30, ns, N))
    print("cholesky payoffs N=%d method=%d" % (N, method), flush=True)
    pay, price, stat = rb.payoff_samples_v2(0.04, o["H"], o["eta"], o["rho"], o["T"], o["K"], N, Z,
                                            method=method, compensator=rb.COMP_EXACT,
                                            scheme=rb.SCHEME_CHOLESKY)
    ref = orc.payoff_chol_samples(0.04, o["H"], o["eta"], o["rho"], o["T"], o["K"], N, Z,
                                  comp=1, method=method)
    assert np.max(np.abs(pay - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-9


def test_cholesky_seeded_prices_match_oracle(rb, orc):
    o = OPTS
    N, M, seed = 64, 3000, [8, 8]
    p, iv, s = rb.price_v2(0.04, o["H"], o["eta"], o["rho"], o["T"], o["K"], N, M, seed,
                           scheme=rb.SCHEME_CHOLESKY)
    Z = orc.normals_philox(seed, 0, M, 2, N)
    pr, st = orc.price_stat(orc.payoff_chol_samples(0.04, o["H"], o["eta"], o["rho"], o["T"],
                                                    o["K"], N, Z))
    assert np.max(np.abs(p - pr)) < 1e-11


def test_reference_recorded_cholesky_and_hybrid_prices(rb):
    args = (0.04, [0.06], [2.5], [-0.8], [0.8], [1.0], 1000, 200000)
    pc, _, sc = rb.price_v2(*args, [12534, 45567], method=rb.METHOD_MC, compensator=rb.COMP_EXACT,
                            scheme=rb.SCHEME_CHOLESKY)
    print("cholesky %.8f +- %.2e" % (pc[0], sc[0]), flush=True)
    assert abs(pc[0] - 0.05784862) <= 4.5 * np.hypot(sc[0], 0.00011394)
    ph, _, sh = rb.price_v2(*args, [53534, 25345], method=rb.METHOD_MC, compensator=rb.COMP_EXACT)
    print("hybrid %.8f +- %.2e" % (ph[0], sh[0]), flush=True)
    assert abs(ph[0] - 0.05761252) <= 4.5 * np.hypot(sh[0], 0.0001136558)


def test_cholesky_scheme_errors(rb):
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.1], [1.5], [-0.7], [1.0], [1.0], 64, 1000, [1],
                    precision=rb.PREC_SPLIT_FP16, scheme=rb.SCHEME_CHOLESKY)
    assert e.value.code == 1
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.5], [1.5], [-0.7], [1.0], [1.0], 64, 1000, [1],
                    scheme=rb.SCHEME_CHOLESKY)
    assert e.value.code == 1
