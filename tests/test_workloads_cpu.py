"""The benchmark's synthetic workloads (bench.py) have the shapes the configurations name: the
256-set batched grid covers every step count, and the serving trace's requests carry 10-200
options with N = max(32, 256 T) over their maturities.  Host only."""
import numpy as np

import bench


def test_256_set_grid_covers_every_step_count():
    from test_gpu_configs import SETS256_ORACLE_SETS
    ds = bench.batched_requests()
    assert len(ds) == 256
    assert sorted({d["N"] for d in ds}) == [32, 64, 128, 256, 512]
    assert {ds[i]["N"] for i in SETS256_ORACLE_SETS} == {32, 64, 128, 256, 512}
    for d in ds:
        assert 0.02 <= d["H"][0] <= 0.45 and 0.5 <= d["eta"][0] <= 4.0 and -0.99 <= d["rho"][0] <= -0.1
        assert 0.01 <= d["xi"] <= 0.16 and 0.05 <= d["T"][0] <= 2.0 and 2 ** 14 - 1 <= d["M"] <= 2 ** 20 + 1
        assert len(d["T"]) == 21 and np.allclose(d["K"], np.linspace(0.8, 1.2, 21))


def test_serving_requests_span_the_grid():
    ds = bench.serving_requests(4096)
    assert len(ds) == 4096
    assert min(d["N"] for d in ds) == 32 and max(d["N"] for d in ds) == 512
    assert min(len(d["T"]) for d in ds) >= 10 and max(len(d["T"]) for d in ds) <= 200
    for d in ds:
        # N = max(32, ceil(256 T)) of the request's longest drawn maturity (which a request's
        # options may not all reach)
        assert d["N"] >= max(32, int(np.ceil(256 * max(d["T"]))))
        assert 2 ** 14 - 1 <= d["M"] <= 2 ** 21 + 1 and len(set(d["T"])) <= 5
    # the slice the GPU test runs is the head of the full trace
    assert [d["seed"] for d in bench.serving_requests(384)] == [d["seed"] for d in ds[:384]]
