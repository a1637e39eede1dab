"""CPU check of the LDS FFT convolution behind the long-path kernels: the group functions and pass
schedules of rb_fft_lds.h, compiled for the host and run pass by pass, against a direct FP64
linear convolution (fft_lds_emulation.cpp) for transform lengths 2^10..2^14 in both transform
orders, with ragged N and the pad slots checked untouched."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "rbergomi_amd", "csrc")


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_fft_lds_schedule_matches_direct_convolution(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "fft_lds_emulation")
    cmd = [hipcc, "-O1", "-std=c++17", "-x", "hip", "--cuda-host-only", "-I" + CSRC,
           os.path.join(HERE, "fft_lds_emulation.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().endswith("OK")
