"""The C++ interface (include/interface.h) called from C++: tests/cpp/interface_test.cpp, built by
__graft_entry__.build() and linked to librbergomi_mi355x.so, calls ComputeIV_singleThread,
ComputeIV_multiThread, ComputeIVRT_singleThread and ComputeIVRT_multiThread on the reference's
inputs as testRBergLib/tests.cpp:13-80 does and checks its known prices within 0.005 + stat, the
implied vols, ResultSimple's echo fields in their order (H, rho, eta, T, K) and equality with
the C entry points on the same seed."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "interface_test")


def test_cpp_interface_program():
    assert os.path.exists(EXE), "tests/cpp/interface_test is not built (run __graft_entry__.build())"
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(p.stdout[-4000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert p.stdout.strip().splitlines()[-1] == "ALL OK"
    assert p.stdout.count("ok  ") >= 4 * (4 * 3 + 5) + 6
