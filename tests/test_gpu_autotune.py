"""GPU checks of the fused kernel's tile shapes (path blocks of 32 paths per tile) and of the
autotuner that picks them.

Every valid shape prices the same paths (same Philox counters) and differs only in the order of
partial sums: FP64 sums for the Romano-Touzi estimator (prices agree to 1e-12), and for plain
Monte Carlo also the FP32 sums over a wave's row blocks (FP32 rounding, prices agree to 1e-8).
All stay within the 1e-4 budget of the FP64 path; after tuning, a request prices identically
alone and in a batch.
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


@pytest.fixture()
def serving(rb):
    from rbergomi_amd import serving
    serving.tune_clear()
    yield serving
    serving.tune_clear()


ARGS = dict(xi=0.04, H=[0.1, 0.3], eta=[1.5, 0.9], rho=[-0.7, -0.3], T=[0.1, 0.5, 1.0], K=[0.95, 1.0, 1.1])


def price(rb, N, M, precision, method=0):
    a = ARGS
    return rb.price_v2(a["xi"], a["H"], a["eta"], a["rho"], a["T"], a["K"], N, M, [9, N],
                       precision=precision, method=method)


@pytest.mark.parametrize("N,shapes", [(32, [1, 2, 4]), (64, [1, 2, 4]), (96, [1, 2, 4]), (160, [1, 2, 4]),
                                      (256, [1, 2])])
@pytest.mark.parametrize("method", [0, 1])
def test_every_tile_shape_prices_alike(rb, serving, tmp_path, N, shapes, method):
    M = 3000
    say("tile shapes N=%d method=%d" % (N, method))
    p64, _, s64 = price(rb, N, M, rb.PREC_FP64, method)
    ref = None
    for pb in shapes:
        f = tmp_path / ("t%d.txt" % pb)
        Np = (N + 31) // 32 * 32
        f.write_text("%d 0 %d %d\n" % (Np, method, pb))
        serving.tune_load(f)
        p, _, s = price(rb, N, M, rb.PREC_SPLIT_FP16, method)
        assert np.max(np.abs(p - p64)) <= 1e-4 and np.max(np.abs(s - s64)) <= 1e-4
        if ref is None:
            ref = p
        say("   PB=%d max |price - first shape| = %.2e" % (pb, np.max(np.abs(p - ref))))
        assert np.max(np.abs(p - ref)) <= (1e-12 if method == 0 else 1e-8)


def test_autotune_records_choices_and_batches_match_alone(rb, serving):
    from rbergomi_amd.serving import Request, autotune, price_batch
    shape = Request(0.04, [0.1], [1.5], [-0.7], [0.25, 1.0], [1.0, 1.1], 0, 1 << 15, [1, 2])
    Ns = [64, 128, 200, 512]
    say("autotune", Ns)
    res = autotune(shape, Ns, precision=rb.PREC_SPLIT_FP16)
    say("   ", res)
    assert set(res) == set(Ns)
    for n, (pb, ms) in res.items():
        assert pb in (1, 2, 4) and ms > 0
    table = {(e[0], e[1], e[2]): e[3] for e in serving.tune_table()}
    for n in Ns:
        assert table[((n + 31) // 32 * 32, 0, 0)] == res[n][0]
    reqs = [Request(0.05, [0.12], [1.4], [-0.6], [0.3, 0.9], [0.9, 1.05], n, 5000, [3, n]) for n in Ns]
    price_batch(reqs, precision=rb.PREC_SPLIT_FP16)
    for r in reqs:
        p, iv, st = rb.price_v2(r.xi, r.H, r.eta, r.rho, r.T, r.K, r.N, r.M, r.seed,
                                precision=rb.PREC_SPLIT_FP16)
        assert np.array_equal(p, r.price) and np.array_equal(st, r.stat)


def test_autotune_rejects_out_of_range(rb, serving):
    from rbergomi_amd.serving import Request, autotune
    shape = Request(0.04, [0.1], [1.5], [-0.7], [1.0], [1.0], 0, 4096, [1])
    with pytest.raises(rb.RbError) as e:
        autotune(shape, [2048])
    assert e.value.code == 5
