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
