"""The oracle's unordered-rows restatement (updatePayoffUnordered, rBergomi/rBergomiMT.cpp:939-960;
per-row xi as ML_console_ST/ParamTot.h:25-43), pinned against the tensor-form restatement, which
the reference's own per-path payoffs (debug8) and closed forms pin: rows laid out as the
reference's zipped tensor must reproduce it, a per-row xi must equal separate scalar-xi runs, and
the rows are independent of each other (permutation)."""
import numpy as np
import pytest


def _tensor_rows(H, eta, rho, T, K):
    rows = [(h, t) for h in range(len(H)) for t in range(len(T))]
    pick = lambda v, idx: [v[i] for i in idx]
    hs, ts = [h for h, _ in rows], [t for _, t in rows]
    return pick(H, hs), pick(eta, hs), pick(rho, hs), pick(T, ts), pick(K, ts)


@pytest.mark.parametrize("comp", [0, 1])
def test_rows_reproduce_the_tensor_form(orc, comp):
    rng = np.random.default_rng(5)
    N = 40
    Z = rng.standard_normal((30, 2, N))
    H, eta, rho = [0.07, 0.3, 0.07], [1.9, 0.9, 1.2], [-0.9, -0.3, -0.5]
    T, K = [0.25, 1.0, 1.0, 2.0], [0.9, 1.0, 1.2, 0.8]
    ref = orc.payoff_rt_samples(0.05, H, eta, rho, T, K, N, Z, comp=comp)
    got = orc.payoff_rows_samples(0.05, *_tensor_rows(H, eta, rho, T, K), N, Z, comp=comp)
    assert np.array_equal(got, ref)


def test_rows_per_row_xi_equals_scalar_runs(orc):
    rng = np.random.default_rng(6)
    N = 32
    Z = rng.standard_normal((25, 2, N))
    xi = [0.01, 0.04, 0.09, 0.16]
    H, eta, rho = [0.1, 0.2, 0.1, 0.4], [1.5, 2.0, 1.0, 0.7], [-0.7, -0.9, -0.2, -0.5]
    T, K = [0.5, 1.0, 2.0, 0.1], [1.0, 0.8, 1.3, 1.0]
    got = orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z)
    for i in range(4):
        one = orc.payoff_rt_samples(xi[i], [H[i]], [eta[i]], [rho[i]], [T[i]], [K[i]], N, Z)
        assert np.array_equal(got[:, i], one[:, 0])


def test_rows_are_independent_of_order(orc):
    rng = np.random.default_rng(7)
    N = 24
    Z = rng.standard_normal((20, 2, N))
    n = 9
    xi = rng.uniform(0.01, 0.1, n)
    H = rng.uniform(0.05, 0.45, n)
    eta, rho = rng.uniform(0.5, 3.0, n), rng.uniform(-0.95, -0.1, n)
    T, K = rng.uniform(0.1, 2.0, n), np.exp(rng.uniform(-0.3, 0.3, n))
    a = orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z)
    perm = rng.permutation(n)
    b = orc.payoff_rows_samples(xi[perm], H[perm], eta[perm], rho[perm], T[perm], K[perm], N, Z)
    assert np.array_equal(a[:, perm], b)


def _turbo_rows_case():
    K = [0.8, 0.95, 1.0, 1.1, 1.3]
    n = len(K)
    return (np.full(n, 0.04), np.full(n, 0.1), np.full(n, 1.5), np.full(n, -0.7),
            np.full(n, 1.0), np.array(K))


def test_turbo_is_unbiased_against_plain_rt(orc):
    """The turbo-charged estimator (control variate + antithetics, ML_console_ST/RBergomi.cpp:
    219-288) and the plain RT estimator estimate the same price: within 4.5 combined SE."""
    xi, H, eta, rho, T, K = _turbo_rows_case()
    N = 24
    rng = np.random.default_rng(11)
    Zc = rng.standard_normal((1000, 2, N))
    Z = rng.standard_normal((3000, 2, N))
    tb = orc.turbo_rows(xi, H, eta, rho, T, K, N, Zc, Z)
    Zp = rng.standard_normal((20000, 2, N))
    pp, sp = orc.price_stat(orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Zp))
    assert np.all(np.abs(tb["price"] - pp) <= 4.5 * np.sqrt(tb["stat"] ** 2 + sp ** 2))
    # the variance reduction is real: same path count, smaller standard error
    _, s_same = orc.price_stat(orc.payoff_rows_samples(xi, H, eta, rho, T, K, N, Z))
    assert np.all(tb["stat"] < s_same)
    assert np.all(tb["Q"] > 0) and np.all(np.isfinite(tb["alpha"]))


def test_turbo_eta_zero_closed_form(orc, ):
    """eta = 0: v = xi, Iv = xi T on every path, so E[X] = C_BS(1, K, T, sqrt(xi)), Q = 2 xi T
    exactly, and X, Y are functions of one lognormal S' (the control variate removes most of
    the variance)."""
    from scipy.stats import norm
    xi, H, eta, rho, T, K = _turbo_rows_case()
    eta = np.zeros_like(eta)
    N = 16
    rng = np.random.default_rng(12)
    tb = orc.turbo_rows(xi, H, eta, rho, T, K, N, rng.standard_normal((500, 2, N)),
                        rng.standard_normal((4000, 2, N)))
    s = np.sqrt(xi * T)
    d1 = -np.log(K) / s + 0.5 * s
    exact = norm.cdf(d1) - K * norm.cdf(d1 - s)
    assert np.allclose(tb["Q"], 2 * xi * T, rtol=1e-12)
    assert np.all(np.abs(tb["price"] - exact) <= 4.5 * tb["stat"] + 1e-12)
    assert np.all(tb["exceptions"] == 0)


def test_turbo_counts_q_exceptions(orc):
    """With a tiny pilot, main paths can exceed Q = max 2 Iv; those evaluations use qmax and are
    counted (ML_console_ST/RBergomi.cpp:486-488), and the price stays finite."""
    xi, H, eta, rho, T, K = _turbo_rows_case()
    eta = np.full_like(eta, 3.0)
    N = 16
    rng = np.random.default_rng(13)
    tb = orc.turbo_rows(xi, H, eta, rho, T, K, N, rng.standard_normal((3, 2, N)),
                        rng.standard_normal((2000, 2, N)), qmax=50.0)
    assert np.all(tb["exceptions"] > 0)
    assert np.all(np.isfinite(tb["price"])) and np.all(np.isfinite(tb["stat"]))
