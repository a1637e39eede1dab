"""ctypes binding of librbergomi_mi355x.so (the HIP product library).

There is no CPU fallback: if the library is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbergomi_mi355x.so")

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
lp = ctypes.POINTER(ctypes.c_long)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)


class RbConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("compensator", ctypes.c_int),
                ("method", ctypes.c_int), ("scheme", ctypes.c_int)]


PREC_FP64, PREC_FP32, PREC_SPLIT_FP16, PREC_SPLIT_BF16, PREC_FP32_GEMM, PREC_FP32_FFT = 0, 1, 2, 3, 4, 5
COMP_REF, COMP_EXACT = 0, 1
METHOD_RT, METHOD_MC = 0, 1
SCHEME_HYBRID, SCHEME_CHOLESKY = 0, 1
STATUS = {0: "RB_OK", 1: "RB_ERR_ARG", 2: "RB_ERR_SIZE", 3: "RB_ERR_DEVICE", 4: "RB_ERR_NOMEM",
          5: "RB_ERR_LIMIT"}

_lib = None


class RbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


def load():
    """Load the HIP library; raises OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("librbergomi_mi355x.so is not built (run __graft_entry__.build()); "
                      "there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    d, i, l, vp = ctypes.c_double, ctypes.c_int, ctypes.c_long, ctypes.c_void_p
    cfgp = ctypes.POINTER(RbConfig)
    legacy = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, i, dp, dp, dp, dp, dp, dp, dp, dp, i]
    L.computeIV_tensor.argtypes = legacy
    L.computeIV_tensor.restype = None
    L.computeIVRT_tensor.argtypes = legacy
    L.computeIVRT_tensor.restype = None
    L.computeIVRT_tensor_quadrature.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, dp, i, dp]
    L.computeIVRT_tensor_quadrature.restype = None
    L.rb_last_error.restype = ctypes.c_char_p
    L.rb_price_v2.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, cfgp, dp, dp, dp, i]
    L.rb_payoff_samples_v2.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, dp, cfgp, dp, dp, dp]
    L.rb_plan_create.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, cfgp, vp,
                                 ctypes.POINTER(vp)]
    L.rb_price_rows.argtypes = [dp, dp, dp, dp, dp, dp, i, i, l, u64p, i, cfgp, dp, dp, dp]
    L.rb_plan_create_rows.argtypes = [dp, dp, dp, dp, dp, dp, i, i, l, u64p, i, cfgp, vp,
                                      ctypes.POINTER(vp)]
    L.rb_price_rows_turbo.argtypes = [dp, dp, dp, dp, dp, dp, i, i, l, u64p, i, l, d, dp, dp, dp,
                                      dp, dp, lp]
    L.rb_payoff_rows_samples.argtypes = [dp, dp, dp, dp, dp, dp, i, i, l, dp, cfgp, dp, dp, dp]
    L.rb_plan_run.argtypes = [vp]
    L.rb_plan_results.argtypes = [vp, dp, dp, dp]
    L.rb_plan_set_profiling.argtypes = [vp, i]
    L.rb_plan_set_graph.argtypes = [vp, i]
    L.rb_plan_kernel_times.argtypes = [vp, ctypes.c_char_p, dp, ip, i, ip]
    L.rb_plan_info.argtypes = [vp, ip, ip, ip, lp]
    L.rb_plan_destroy.argtypes = [vp]
    L.rb_plan_destroy.restype = None
    L.rb_layout.argtypes = [dp, dp, dp, i, dp, dp, i, i, i, ip, ip, ip, ip]
    L.rb_probe_philox.argtypes = [u32p, u32p, l, u32p]
    L.rb_probe_normals.argtypes = [u64p, i, l, l, i, i, dp]
    L.rb_probe_wtilde.argtypes = [dp, l, i, d, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.rb_probe_mfma_layout.argtypes = [fp, fp, i, fp]
    L.rb_probe_payoff.argtypes = [dp, dp, dp, i, dp, fp, fp, i]
    L.rb_probe_copy.argtypes = [l, i, dp]
    L.rb_probe_issue.argtypes = [dp, ctypes.POINTER(ctypes.c_int)]
    L.rb_probe_chol.argtypes = [d, i, dp, dp, dp]
    L.rb_probe_sobol.argtypes = [i, l, l, dp]
    L.rb_probe_norm_inv.argtypes = [dp, l, dp]
    L.rb_price_sobol.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, cfgp, dp, dp, dp, i]
    L.rb_price_sobol_bridge.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, cfgp, dp, dp, dp, i]
    L.rb_probe_bridge.argtypes = [dp, l, i, dp]
    L.rb_probe_integrals.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, dp, cfgp, dp]
    L.rb_iv_reference.argtypes = [dp, dp, dp, i, dp]
    L.rb_price_noalloc.argtypes = [d, dp, dp, dp, i, dp, dp, i, i, l, u64p, i, cfgp, vp, dp, dp, dp, i]
    _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RbError(rc, load().rb_last_error().decode())
