"""The library inside a process that uses torch on the GPU.

torch ships its own HIP runtime (soname libamdhip64.so.7, as the system one).  When torch is
imported first, the dynamic linker resolves this library's HIP dependency to that already
loaded runtime, so both share one runtime and a torch stream handle can be given to a plan.
The check runs in a fresh process (the load order is per process) and compares its prices with
the same call made here.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import torch
x = torch.randn(1 << 16, device="cuda")
s = torch.cuda.Stream()
import numpy as np
from rbergomi_amd import cinterface as ci
maps = open("/proc/self/maps").read()
hip = sorted({l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l})
p = ci.Plan(0.04, [0.1], [1.5], [-0.7], [0.5, 1.0], [1.0, 0.9], 128, 1 << 14, [7],
            precision=2, stream=s.cuda_stream, graph=True)
p.run()
p.run()
price = p.results()[0]
p.close()
y = (x * 2).sum().item()
print(json.dumps({"hip": hip, "price": price.tolist(), "torch_ok": bool(np.isfinite(y))}))
"""


def test_torch_first_shares_one_runtime():
    from rbergomi_amd import cinterface as ci
    print("torch interop child", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert len(out["hip"]) == 1, out["hip"]
    assert out["torch_ok"]
    ref = ci.price_v2(0.04, [0.1], [1.5], [-0.7], [0.5, 1.0], [1.0, 0.9], 128, 1 << 14, [7],
                      precision=2)[0]
    assert np.array_equal(np.array(out["price"]), ref)
