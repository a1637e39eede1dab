"""rbergomi_amd -- MI355X-native rough-Bergomi Monte Carlo pricer (HIP, gfx950).

The product is the C-ABI library librbergomi_mi355x.so (include/rbergomi_mi355x.h,
include/interface.h); this package builds it and mirrors the reference's pricing interface
in Python through ctypes.
"""
from ._lib import (COMP_EXACT, COMP_REF, METHOD_MC, METHOD_RT, PREC_FP32, PREC_FP32_FFT, PREC_FP32_GEMM, PREC_FP64,
                   PREC_SPLIT_BF16, PREC_SPLIT_FP16, SCHEME_CHOLESKY, SCHEME_HYBRID, RbError,
                   load)
from .cinterface import (Plan, bridge_increments, iv_reference, computeIV, computeIV_RT, computeIVRT_tensor_quadrature, layout,
                         payoff_rows_samples, payoff_samples_v2, price_noalloc, price_rows,
                         price_rows_turbo, price_sobol, price_v2, sobol_points)

__all__ = ["iv_reference", "computeIV", "computeIV_RT", "computeIVRT_tensor_quadrature", "price_v2",
           "payoff_samples_v2", "price_noalloc", "price_rows", "price_rows_turbo", "price_sobol", "sobol_points", "bridge_increments", "payoff_rows_samples", "Plan", "layout", "load", "RbError", "PREC_FP64", "PREC_FP32", "PREC_FP32_GEMM", "PREC_FP32_FFT",
           "PREC_SPLIT_FP16", "PREC_SPLIT_BF16", "COMP_REF", "COMP_EXACT", "METHOD_RT",
           "METHOD_MC", "SCHEME_HYBRID", "SCHEME_CHOLESKY"]
