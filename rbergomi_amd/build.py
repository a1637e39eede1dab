"""Build recipe of librbergomi_mi355x.so (hipcc, gfx950) -- used by __graft_entry__.build().

hipcc compiles the HIP kernels and the C++ host code into one shared library in-tree
(rbergomi_amd/librbergomi_mi355x.so), so that it travels with the repository snapshot.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librbergomi_mi355x.so")

SOURCES = ["rb_staged.hip", "rb_fused.hip", "rb_fused_rt6.hip", "rb_fused_rt2.hip", "rb_fused_rt1.hip", "rb_fused_rt5.hip", "rb_fused_rt4.hip", "rb_fused_rt3.hip", "rb_fused_mc.hip", "rb_fused_mc8.hip", "rb_fused_mc2.hip", "rb_fused_mc1.hip", "rb_fused_mc5.hip", "rb_fused_mc4.hip", "rb_fused_mc3.hip", "rb_fft.hip",
           "rb_turbo.hip", "rb_chol.hip", "rb_probe.hip", "rb_volterra64.hip", "rb_fused_fft.hip", "rb_engine.cpp", "rb_chol.cpp", "rb_sobol.cpp",
           "rb_fused_host.cpp", "rb_capi.cpp", "rb_server.cpp", "rb_noalloc.hip"]
# Packed FP32 instructions are disabled for all device code (see the package notes): the
# compiler would otherwise form v_pk_* from ordinary float code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Wall",
          "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include")]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=4):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"]
        cmd = [hipcc()] + COMMON + DEVICE_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 target feature: harmless, filtered
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if verbose or p.returncode:
            sys.stderr.write("\n".join(lines) + ("\n" if lines else ""))
        if p.returncode:
            sys.stderr.write("FAILED: " + " ".join(cmd) + "\n")
            failed = True
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    link = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.check_call(link)
    os.replace(tmp, LIB)
    return LIB


def build_interface_test():
    """The C++ caller of include/interface.h (tests/cpp/interface_test.cpp), compiled with g++ and
    linked to the library; the GPU test runs it."""
    root = os.path.abspath(os.path.join(HERE, ".."))
    src = os.path.join(root, "tests", "cpp", "interface_test.cpp")
    out = os.path.join(root, "tests", "cpp", "interface_test")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(root, "include"), src, "-o", out,
           "-L" + HERE, "-lrbergomi_mi355x", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../rbergomi_amd"]
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
