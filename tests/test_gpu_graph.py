"""GPU checks of graph replay (rb_plan_set_graph).

A plan in graph mode records its launches into a hipGraph on the first run and replays that
graph on every later run.  Each replay must give exactly the eager run's prices, implied vols
and standard errors: the work-queue counter is cleared by a kernel inside the graph (no memset
node), so a replay that ran with stale device state would show up as a changed result.  Between
replays the test runs other pricing calls eagerly (on their own plans and streams) and a device
copy kernel, so the graph's buffers see unrelated traffic in between.
"""
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


# (label, precision, N, M, method, scheme)
CASES = [
    ("fused_fp16", 2, 256, 1 << 15, 0, 0),
    ("fused_bf16", 3, 128, 20000, 0, 0),
    ("fused_fp32", 4, 512, 9000, 0, 0),
    ("fused_fft_fp16", 2, 2048, 3001, 0, 0),
    ("fused_fp16_mc", 2, 96, 20000, 1, 0),
    ("staged_fp64", 0, 100, 1 << 14, 0, 0),
    ("staged_fp64_mc", 0, 64, 1 << 14, 1, 0),
    ("fft_fp32", 1, 1100, 4096, 0, 0),
    ("cholesky_fp64", 0, 80, 8192, 0, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_match_eager(rb, case):
    import ctypes
    from rbergomi_amd import cinterface as ci
    label, prec, N, M, method, scheme = case
    H, eta, rho = [0.07, 0.1], [1.9, 1.5], [-0.9, -0.7]
    T = [0.25, 1.0, 1.0, 2.0]
    K = [0.9, 1.0, 1.2, 0.8]
    say("graph case", label, "N", N, "M", M)
    eager = ci.Plan(0.04, H, eta, rho, T, K, N, M, [11, 5], precision=prec, method=method,
                    scheme=scheme)
    eager.run()
    ref = eager.results()
    eager.close()
    assert np.all(np.isfinite(ref[0]))

    g = ci.Plan(0.04, H, eta, rho, T, K, N, M, [11, 5], precision=prec, method=method, graph=True,
                scheme=scheme)
    for rep in range(4):
        g.run()
        got = g.results()
        for a, b, name in zip(got, ref, ("price", "iv", "stat")):
            assert np.array_equal(a, b), (label, rep, name, np.max(np.abs(a - b)))
        # unrelated eager work between replays: another plan on its own stream, and a copy kernel
        other = ci.price_v2(0.04, [0.2], [1.0], [-0.5], [0.5], [1.0], 64, 4096, [rep],
                            precision=prec)
        assert np.all(np.isfinite(other[0]))
        ms = ctypes.c_double()
        assert rb.load().rb_probe_copy(64 << 20, 2, ctypes.byref(ms)) == 0
    g.close()


def _batch_requests(seed, N_list=(32, 64, 200, 512, 96), long_N=None):
    from rbergomi_amd.serving import Request
    rng = np.random.default_rng(seed)
    out = []
    for i, N in enumerate(N_list):
        T = list(np.repeat(rng.uniform(0.1, 2.0, 2), 3))
        K = list(np.exp(rng.uniform(-0.3, 0.2, 6)))
        out.append(Request(rng.uniform(0.02, 0.1), [rng.uniform(0.05, 0.4)], [rng.uniform(0.8, 3.0)],
                           [rng.uniform(-0.95, -0.2)], T, K, N, 3000 + 517 * i, [seed, i]))
    if long_N:
        out.append(Request(0.04, [0.1], [1.5], [-0.7], [0.5, 1.0], [1.0, 0.9], long_N, 2001, [seed, 99]))
    return out


@pytest.mark.parametrize("long_N", [None, 2048])
def test_batch_graph_replays_and_updates(rb, long_N):
    """rb_batch with graph replay: the replays equal rb_price_batch of the same requests, with other
    pricing calls between them; after rb_batch_update with new parameters of the same shapes, the
    replays equal rb_price_batch of the new requests."""
    from rbergomi_amd.serving import Batch, price_batch
    reqs = _batch_requests(5, long_N=long_N)
    ref = price_batch(_batch_requests(5, long_N=long_N), precision=rb.PREC_SPLIT_FP16)
    with Batch(reqs, precision=rb.PREC_SPLIT_FP16, graph=True) as b:
        for rep in range(4):
            b.run()
            for r, q in zip(reqs, ref):
                assert np.array_equal(r.price, q.price) and np.array_equal(r.stat, q.stat), rep
                assert np.array_equal(r.iv, q.iv, equal_nan=True)
                r.price[:] = 0.0
            rb.price_v2(0.04, [0.1], [1.5], [-0.7], [1.0], [1.0], 64, 1000, [7])  # other work between
        new = _batch_requests(6, long_N=long_N)
        b.update(new)
        ref2 = price_batch(_batch_requests(6, long_N=long_N), precision=rb.PREC_SPLIT_FP16)
        for rep in range(3):
            b.run()
            for r, q in zip(new, ref2):
                assert np.array_equal(r.price, q.price) and np.array_equal(r.stat, q.stat), rep
            assert not np.array_equal(new[0].price, ref[0].price)


def test_batch_update_rejects_changed_layout(rb):
    from rbergomi_amd.serving import Batch, Request
    reqs = _batch_requests(7)
    with Batch(reqs, precision=rb.PREC_SPLIT_BF16) as b:
        b.run()
        bad = _batch_requests(8)
        bad[1] = Request(0.05, [0.1, 0.2], [1.5, 1.5], [-0.7, -0.7], bad[1].T[:3], bad[1].K[:3], bad[1].N, bad[1].M, [1])
        with pytest.raises(rb.RbError) as e:
            b.update(bad)
        assert e.value.code == 1
        b.run()  # the batch still prices its old requests
