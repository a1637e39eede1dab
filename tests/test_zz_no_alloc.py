"""rb_price_noalloc (pricing without device allocation) against rb_price_v2, through the C ABI.

The allocation-free entry point runs the same fused-kernel launches as rb_price_v2 with every
device buffer carved from the library's static per-device scratch, so its prices, implied vols
and standard errors must equal rb_price_v2's bit for bit.  Two calls in a row, the second
smaller, check that the scratch keeps no state that leaks from one call into the next (the
work-queue counter is cleared by a kernel, every other buffer is written before it is read).
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


def _case_big():
    # five maturities x 41 strikes at N = 512 (the single-set workload's option grid), M = 2^17
    T = np.repeat([0.1, 0.25, 0.5, 1.0, 2.0], 41)
    K = np.exp(np.tile(np.linspace(-0.4, 0.2, 41), 5))
    return dict(xi=0.04, H=[0.1], eta=[1.5], rho=[-0.7], T=T, K=K, N=512, M=2 ** 17, seed=[9, 1])


def _case_small():
    # a smaller second call: two parameter sets, three options, N = 64, odd M
    return dict(xi=0.055, H=[0.07, 0.2], eta=[1.9, 1.1], rho=[-0.9, -0.4], T=[0.5, 1.0, 1.0],
                K=[1.0, 0.9, 1.2], N=64, M=3001, seed=[9, 2])


def _args(c):
    return (c["xi"], c["H"], c["eta"], c["rho"], c["T"], c["K"], c["N"], c["M"], c["seed"])


@pytest.mark.parametrize("prec,method", [("PREC_SPLIT_FP16", "METHOD_RT"), ("PREC_SPLIT_BF16", "METHOD_RT"),
                                         ("PREC_FP32_GEMM", "METHOD_RT"), ("PREC_SPLIT_FP16", "METHOD_MC")])
def test_noalloc_equals_price_v2_twice(rb, prec, method):
    kw = dict(precision=getattr(rb, prec), method=getattr(rb, method))
    for c in (_case_big(), _case_small()):
        say("noalloc %s %s N=%d M=%d options=%d" % (prec, method, c["N"], c["M"], len(c["H"]) * len(c["T"])))
        want = rb.price_v2(*_args(c), **kw)
        got = rb.price_noalloc(*_args(c), **kw)
        for w, g in zip(want, got):
            assert np.array_equal(w, g, equal_nan=True)


def test_noalloc_rejects_what_it_cannot_hold(rb):
    c = _case_big()
    args = list(_args(c))
    with pytest.raises(rb.RbError) as e:  # per-item option sums alone: 205 x 2^21 x 16 B = 6.7 GB
        rb.price_noalloc(*args[:7], 2 ** 31, args[8])
    assert e.value.code == 5 and "scratch" in str(e.value)
    with pytest.raises(rb.RbError) as e:  # the FP64 path allocates its staging buffers
        rb.price_noalloc(*args, precision=rb.PREC_FP64)
    assert e.value.code == 1
    with pytest.raises(rb.RbError) as e:  # long paths run the FFT kernel
        rb.price_noalloc(*args[:6], 2048, 1000, args[8])
    assert e.value.code == 5
    # and a valid call afterwards still equals rb_price_v2
    s = _case_small()
    for w, g in zip(rb.price_v2(*_args(s), precision=rb.PREC_SPLIT_FP16),
                    rb.price_noalloc(*_args(s), precision=rb.PREC_SPLIT_FP16)):
        assert np.array_equal(w, g, equal_nan=True)
