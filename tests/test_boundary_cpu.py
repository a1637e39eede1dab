"""CPU-only checks of the C-ABI library: it loads, exports every declared entry point, plans
work on the host, and keeps the reference's argument-error behaviour (no device needed)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rbergomi_amd", "librbergomi_mi355x.so")


def declared_c_functions():
    with open(os.path.join(ROOT, "include", "rbergomi_mi355x.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", text)) -
                  {"if", "sizeof", "return", "defined"})


def test_library_loads_and_exports_c_abi():
    lib = ctypes.CDLL(LIB)
    names = declared_c_functions()
    assert {"computeIV_tensor", "computeIVRT_tensor", "computeIVRT_tensor_quadrature"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r" T %s$" % n, nm, re.M), n


def test_library_exports_cpp_interface():
    nm = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True).stdout
    with open(os.path.join(ROOT, "include", "interface.h")) as f:
        decl = re.findall(r"^ResultSimple (\w+)\(", f.read(), re.M)
    assert len(decl) == 4
    for n in decl:
        assert re.search(r"\b%s\(double, std::vector<double" % n, nm), n


def test_layout_dedupes_units_groups():
    import rbergomi_amd as rb
    H = [0.07, 0.1, 0.07]
    eta = [2.2, 2.4, 2.2]
    rho = [-0.9, -0.85, -0.5]
    T = [0.5, 1.0, 0.5]
    K = [0.8, 1.0, 1.2]
    nu, ng, gofo, uofg = rb.layout(H, eta, rho, T, K, 256)
    assert nu == 2 and ng == 4
    # I = h*op_size + t; triples 0 and 2 share (H, eta) -> same groups; T 0.5 repeats
    assert list(gofo) == [0, 1, 0, 2, 3, 2, 0, 1, 0]
    assert list(uofg) == [0, 0, 1, 1]
    # plain MC separates rho
    nu, ng, gofo, _ = rb.layout(H, eta, rho, T, K, 256, method=rb.METHOD_MC)
    assert nu == 2 and ng == 6 and gofo[6] != gofo[0]


def _legacy_script(body):
    return ("import sys, numpy as np; sys.path.insert(0, %r)\n"
            "import rbergomi_amd._lib as L, ctypes\n"
            "lib = L.load(); dp = L.dp\n"
            "a = lambda x: np.ascontiguousarray(x, dtype=np.float64)\n"
            "H, e, r, T, K = a([0.1]), a([1.5]), a([-0.7]), a([1.0, 2.0]), a([1.0, 1.1])\n"
            "seed = np.array([1, 2], dtype=np.uint64)\n"
            "o = [np.zeros(2) for _ in range(8)]\n"
            "P = lambda x: x.ctypes.data_as(dp)\n" % ROOT) + body


@pytest.mark.parametrize("out_size,threads,code", [(3, 4, 17), (2, 0, 21)])
def test_legacy_argument_errors_exit_like_reference(out_size, threads, code):
    # rBergLib/cinterface.cpp:54-59 (exit 17) and :77-81 (exit 21), checked before any device work
    body = ("lib.computeIVRT_tensor(0.04, P(H), P(e), P(r), 1, P(T), P(K), 2, 16, 100, "
            "seed.ctypes.data_as(L.u64p), 2, %d, *[P(x) for x in o], %d)\n" % (threads, out_size))
    p = subprocess.run([sys.executable, "-c", _legacy_script(body)], capture_output=True, text=True)
    assert p.returncode == code, (p.returncode, p.stderr)
    assert ("wrong size" in p.stderr) if code == 17 else ("positive integer" in p.stderr)


def test_legacy_invalid_values_give_nan_like_reference():
    # H = 0: the reference divides by sqrt(2H) and returns NaN; no device work is started
    body = ("H[0] = 0.0\n"
            "lib.computeIVRT_tensor(0.04, P(H), P(e), P(r), 1, P(T), P(K), 2, 16, 100, "
            "seed.ctypes.data_as(L.u64p), 2, 1, *[P(x) for x in o], 2)\n"
            "assert np.all(np.isnan(o[0])) and np.all(np.isnan(o[2]))\n"
            "assert list(o[3]) == [0.0, 0.0] and list(o[6]) == [1.0, 2.0]\n")
    p = subprocess.run([sys.executable, "-c", _legacy_script(body)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    assert "results set to NaN" in p.stderr


def test_v2_reports_argument_errors():
    import rbergomi_amd as rb
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.0], [1.5], [-0.7], [1.0], [1.0], 16, 100, [1])
    assert e.value.code == 1 and "H must lie" in str(e.value)
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.1], [1.5], [-1.5], [1.0], [1.0], 16, 100, [1])
    assert e.value.code == 1
    with pytest.raises(rb.RbError) as e:
        rb.price_v2(0.04, [0.1], [1.5], [-0.5], [1.0], [1.0], 100000, 100, [1])
    assert e.value.code == 5
    L = rb.load()
    a = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    H, z = a([0.1]), np.zeros(4)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.rb_price_v2(0.04, H.ctypes.data_as(dp), H.ctypes.data_as(dp), H.ctypes.data_as(dp), 1,
                       H.ctypes.data_as(dp), H.ctypes.data_as(dp), 1, 8, 10, None, 0, None,
                       z.ctypes.data_as(dp), z.ctypes.data_as(dp), z.ctypes.data_as(dp), 3)
    assert rc == 2


def test_product_package_does_not_import_oracle():
    import rbergomi_amd
    pkg = os.path.dirname(rbergomi_amd.__file__)
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".h")):
                with open(os.path.join(dirpath, fn)) as f:
                    txt = f.read()
                assert "pyoracle" not in txt and "liborbergomi" not in txt and "orc_" not in txt, fn
