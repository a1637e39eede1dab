"""GPU checks of batched pricing and of the continuous-batching server: every request's
results are bit-identical alone (rb_price_v2) and in a batch or a server mix; admission under
a device-byte budget; oversize requests rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


def make_requests(n, seed=0):
    from rbergomi_amd.serving import Request
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        npar = int(rng.integers(1, 3))
        H = rng.uniform(0.02, 0.45, npar)
        eta = rng.uniform(0.5, 4.0, npar)
        rho = rng.uniform(-0.99, -0.1, npar)
        nopt = int(rng.integers(3, 40))
        T = rng.choice(rng.uniform(0.05, 2.0, 3), nopt)
        K = np.exp(rng.uniform(-0.4, 0.2, nopt))
        N = int(rng.choice([32, 50, 64, 128, 200, 256, 512]))
        M = int(rng.integers(500, 20000))
        reqs.append(Request(rng.uniform(0.01, 0.16), H, eta, rho, T, K, N, M, [i, 77]))
    return reqs


def alone(rb, r, prec):
    return rb.price_v2(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed, precision=prec)


@pytest.mark.parametrize("prec_name", ["fp16", "bf16", "fp32"])
def test_batch_equals_alone(rb, prec_name):
    from rbergomi_amd.serving import price_batch
    prec = {"fp16": rb.PREC_SPLIT_FP16, "bf16": rb.PREC_SPLIT_BF16, "fp32": rb.PREC_FP32_GEMM}[prec_name]
    reqs = make_requests(12, seed=1)
    price_batch(reqs, precision=prec)
    for r in reqs:
        p, iv, st = alone(rb, r, prec)
        assert np.array_equal(p, r.price) and np.array_equal(st, r.stat)
        assert np.array_equal(iv, r.iv, equal_nan=True)


def test_plain_mc_batch_equals_alone(rb):
    # computeIV_tensor's estimator in one batched launch (exact compensator too)
    from rbergomi_amd.serving import price_batch
    reqs = make_requests(8, seed=4)
    price_batch(reqs, precision=rb.PREC_SPLIT_FP16, method=rb.METHOD_MC, compensator=rb.COMP_EXACT)
    for r in reqs:
        p, iv, st = rb.price_v2(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed,
                                precision=rb.PREC_SPLIT_FP16, method=rb.METHOD_MC,
                                compensator=rb.COMP_EXACT)
        assert np.array_equal(p, r.price) and np.array_equal(st, r.stat)
        p64 = rb.price_v2(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed,
                          method=rb.METHOD_MC, compensator=rb.COMP_EXACT)[0]
        assert np.max(np.abs(p - p64)) <= 1e-4


def test_server_mix_equals_alone(rb):
    from rbergomi_amd.serving import Server
    reqs = make_requests(24, seed=2)
    with Server(max_inflight=2, max_batch_requests=5) as srv:
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            srv.wait(r)
            assert r.latency_ms > 0
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 5
    assert st["p99_ms"] >= st["p50_ms"] > 0
    for r in reqs:
        p, iv, s = alone(rb, r, rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price) and np.array_equal(s, r.stat)


def test_server_admits_under_tight_budget(rb):
    from rbergomi_amd.serving import Server
    reqs = make_requests(6, seed=3)
    # budget below two requests: batches are admitted one at a time as memory frees up
    with Server(max_inflight=3, max_batch_requests=8, max_device_bytes=3 << 20) as srv:
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            srv.wait(r)
        st = srv.stats()
    assert st["completed"] == 6
    for r in reqs:
        p, _, s = alone(rb, r, rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price)


def test_server_rejects_oversize_request(rb):
    from rbergomi_amd.serving import Request, Server
    big = Request(0.04, [0.1], [1.5], [-0.7], np.ones(200), np.linspace(0.8, 1.2, 200), 256, 1 << 22, [1])
    with Server(max_device_bytes=1 << 20) as srv:
        with pytest.raises(rb.RbError) as e:
            srv.submit(big)
    assert e.value.code == 5


def test_long_path_requests_in_batches_and_server(rb):
    # requests beyond N = 1024 run the long-path kernel in a launch of their own, between the
    # batched short ones; every result equals pricing the request alone
    from rbergomi_amd.serving import Request, Server, price_batch
    reqs = make_requests(6, seed=9)
    reqs.insert(2, Request(0.05, [0.1, 0.2], [1.5, 1.1], [-0.7, -0.5], [0.5, 1.0], [1.0, 1.1], 2048, 3001, [2, 2048]))
    reqs.append(Request(0.03, [0.05], [2.3], [-0.9], [1.0], [0.9], 1500, 777, [3, 1500]))
    price_batch(reqs, precision=rb.PREC_SPLIT_FP16)
    for r in reqs:
        p, iv, st = alone(rb, r, rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price) and np.array_equal(st, r.stat)
    fresh = [Request(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed) for r in reqs]
    with Server(rb.PREC_SPLIT_FP16, max_inflight=2, max_batch_requests=4) as srv:
        for r in fresh:
            srv.submit(r)
        for r in fresh:
            srv.wait(r)
    for r, q in zip(fresh, reqs):
        assert np.array_equal(r.price, q.price) and np.array_equal(r.stat, q.stat)


def _lds_heavy_pair(method_mc=False):
    """Two requests that each fit one workgroup's LDS (160 KB) alone but not together: N = 1024
    with one maturity (a 64 KB FP16 W1 tile; 69 KB in all) and N = 32 with 8 maturities x 3 rho
    (96 KB of per-path records at four path blocks per tile; 123 KB in all); together 181 KB."""
    from rbergomi_amd.serving import Request
    big_tile = Request(0.04, [0.1], [1.5], [-0.7], [1.0, 1.0], [0.9, 1.1], 1024, 3000, [31, 1])
    mats = np.repeat(np.linspace(0.1, 2.0, 8), 2)
    many_pg = Request(0.05, [0.2] * 3, [1.2] * 3, [-0.8, -0.55, -0.3], mats, np.tile([0.9, 1.1], 8), 32, 5000,
                      [31, 2])
    return [big_tile, many_pg]


def test_lds_incompatible_requests_batch_and_serve(rb):
    # rb_price_batch splits them into launches that fit; the server closes a batch before the
    # request that would not fit with it.  Each result equals pricing the request alone.
    from rbergomi_amd.serving import Request, Server, price_batch
    reqs = _lds_heavy_pair()
    price_batch(reqs, precision=rb.PREC_SPLIT_FP16)
    for r in reqs:
        p, iv, st = alone(rb, r, rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price) and np.array_equal(st, r.stat)
    fresh = [Request(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed) for r in reqs * 3]
    with Server(rb.PREC_SPLIT_FP16, max_inflight=2, max_batch_requests=8) as srv:
        for r in fresh:
            srv.submit(r)
        for r in fresh:
            srv.wait(r)
        st = srv.stats()
    assert st["batches"] >= 2
    for r, q in zip(fresh, reqs * 3):
        assert np.array_equal(r.price, q.price) and np.array_equal(r.stat, q.stat)


def test_chunked_and_unchunked_partials_price_alike(rb):
    # plain MC, N = 1024, five maturities: alone, the cross-wave partials are summed in chunks of
    # three groups (two workgroups per CU fit the LDS beside the 64 KB FP16 W1 tile); in a batch
    # whose other request brings 32 KB of per-path records, the launch is above half a CU's LDS
    # anyway and the five groups go in one chunk.  The sums and their order are the same, so the
    # bits are.
    from rbergomi_amd.serving import Request, price_batch
    T = np.repeat([0.1, 0.25, 0.5, 1.0, 2.0], 3)
    K = np.tile([0.9, 1.0, 1.1], 5)
    a = Request(0.04, [0.1], [1.5], [-0.7], T, K, 1024, 6000, [44, 5])
    mats = np.repeat(np.linspace(0.1, 2.0, 8), 2)
    b = Request(0.05, [0.2, 0.2], [1.2, 1.2], [-0.8, -0.3], mats, np.tile([0.9, 1.1], 8), 32, 5000, [44, 6])
    kw = dict(precision=rb.PREC_SPLIT_FP16, method=rb.METHOD_MC)
    p_alone, _, s_alone = rb.price_v2(a.xi, a.H, a.eta, a.rho, a.T, a.K, a.N, a.M, a.seed, **kw)
    price_batch([a, b], **kw)
    assert np.array_equal(p_alone, a.price) and np.array_equal(s_alone, a.stat)


def test_plain_mc_server_mix_equals_alone(rb):
    from rbergomi_amd.serving import Server
    reqs = make_requests(10, seed=12)
    with Server(rb.PREC_SPLIT_BF16, max_inflight=3, max_batch_requests=4, method=rb.METHOD_MC,
                compensator=rb.COMP_EXACT) as srv:
        srv.set_timing(True)
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            srv.wait(r)
        rep = srv.report()
    assert rep["completed"] == 10 and rep["kernel_launches"] == rep["batches"]
    for r in reqs:
        p, _, s = rb.price_v2(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed, precision=rb.PREC_SPLIT_BF16,
                              method=rb.METHOD_MC, compensator=rb.COMP_EXACT)
        assert np.array_equal(p, r.price) and np.array_equal(s, r.stat)


def test_server_rejects_request_beyond_one_workgroup_lds(rb):
    # 8 maturities x 16 rho at N = 1024: 128 payoff groups of per-path records (128 KB at 32
    # paths per tile) and the 64 KB FP16 W1 tile cannot share one workgroup's 160 KB of LDS, so
    # it is refused at submit with the reason
    from rbergomi_amd.serving import Request, Server
    mats = np.linspace(0.1, 2.0, 8)
    H = [0.1] * 16
    rho = list(np.linspace(-0.9, -0.2, 16))
    big = Request(0.04, H, [1.5] * 16, rho, mats, np.ones(8), 1024, 1000, [1])
    with Server() as srv:
        with pytest.raises(rb.RbError) as e:
            srv.submit(big)
    assert e.value.code == 5 and "LDS" in str(e.value)
