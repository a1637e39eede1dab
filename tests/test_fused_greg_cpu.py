"""CPU check of the fused kernel's groups-per-pass rule (fused_greg in rb_fused.h), compiled for
the host: for every unit size the chosen count is one of the instantiated kernels (1 to 6, or 8),
needs no more passes than eight groups per pass would, and no smaller instantiated count reaches
that many passes (fewer groups per pass leave fewer dead slots)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "rbergomi_amd", "csrc")

SRC = r"""
#include <cstdio>
#include "rb_fused.h"
int main() {
    for (int g = 1; g <= 64; ++g) std::printf("%d %d\n", g, rb::fused_greg(g));
    return 0;
}
"""


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_groups_per_pass_rule(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    src = tmp_path / "greg.cpp"
    src.write_text(SRC)
    exe = str(tmp_path / "greg")
    r = subprocess.run([hipcc, "-O1", "-std=c++17", "-x", "hip", "--cuda-host-only", "-I" + CSRC,
                        str(src), "-o", exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    instantiated = {1, 2, 3, 4, 5, 6, 8}
    passes = lambda g, per: -(-g // per)
    seen = 0
    for line in out.stdout.split("\n"):
        if not line.strip():
            continue
        g, per = map(int, line.split())
        seen += 1
        assert per in instantiated, (g, per)
        assert passes(g, per) == passes(g, 8), (g, per)
        for smaller in sorted(instantiated):
            if smaller >= per:
                break
            assert passes(g, smaller) > passes(g, per), (g, per, smaller)
        if g <= 6:
            assert per == g  # one pass with no dead slot
    assert seen == 64
