"""Python mirror of the reference's pricing interface, on top of the HIP library.

Names and argument meaning follow the reference's R binding of the C ABI
(rBergomiR/RBergomi/src/rbergomi.cpp:16-77, exported as computeIV / computeIV_RT) and the
C interface itself (rBergLib/cinterface.h:62-76).  Every call goes through the C ABI of
librbergomi_mi355x.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import RbConfig, check, dp, u64p


def _arr(x):
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    return a, a.ctypes.data_as(dp)


def _seed(seed):
    s = np.ascontiguousarray(np.atleast_1d(np.asarray(seed, dtype=np.uint64)))
    return s, s.ctypes.data_as(u64p)


def _legacy(fn, xi, H, eta, rho, T, K, N, M, seed, num_threads):
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    s, ps = _seed(seed)
    par_size, op_size = len(H), len(T)
    out_size = par_size * op_size
    outs = {k: np.zeros(out_size) for k in ("price", "iv", "stat", "H", "eta", "rho", "T", "K")}
    p = {k: v.ctypes.data_as(dp) for k, v in outs.items()}
    fn(float(xi), pH, pe, pr, par_size, pT, pK, op_size, int(N), int(M), ps, len(s),
       int(num_threads), p["price"], p["iv"], p["stat"], p["H"], p["eta"], p["rho"], p["T"], p["K"],
       out_size)
    # DataFrame columns of rbergomi.cpp:40-44
    res = {"price": outs["price"], "iv": outs["iv"], "H": outs["H"], "eta": outs["eta"],
           "rho": outs["rho"], "T": outs["T"], "K": outs["K"], "xi": float(xi),
           "stat": outs["stat"], "N": int(N), "M": int(M), "num.threads": int(num_threads)}
    return res


def computeIV(xi, H, eta, rho, T, K, N, M, seed, num_threads=1):
    """Plain Monte Carlo through computeIV_tensor (rbergomi.cpp:17-45).  As in the reference,
    the returned 'eta' column holds rho and the 'rho' column holds eta."""
    return _legacy(_lib.load().computeIV_tensor, xi, H, eta, rho, T, K, N, M, seed, num_threads)


def computeIV_RT(xi, H, eta, rho, T, K, N, M, seed, num_threads=1):
    """Romano-Touzi prices and implied vols through computeIVRT_tensor (rbergomi.cpp:48-77)."""
    return _legacy(_lib.load().computeIVRT_tensor, xi, H, eta, rho, T, K, N, M, seed, num_threads)


def computeIVRT_tensor_quadrature(xi, H, eta, rho, T, K, N, Z, num_threads=1):
    """Per-path RT payoffs from user Gaussians Z (2*N*M values, path-major: W1 block then
    W1perp block) -> payoff [M, out_size]  (rBergLib/cinterface.cpp:102-131)."""
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    Z, pZ = _arr(np.asarray(Z, dtype=np.float64).ravel())
    M = Z.size // (2 * N)
    out = np.zeros((M, len(H) * len(T)))
    _lib.load().computeIVRT_tensor_quadrature(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N),
                                              int(M), pZ, int(num_threads), out.ctypes.data_as(dp))
    return out


def _cfg(precision=_lib.PREC_FP64, compensator=_lib.COMP_REF, method=_lib.METHOD_RT, scheme=0):
    return RbConfig(precision, compensator, method, scheme)


def price_v2(xi, H, eta, rho, T, K, N, M, seed, precision=_lib.PREC_FP64,
             compensator=_lib.COMP_REF, method=_lib.METHOD_RT, scheme=0):
    """Status-checked pricing (rb_price_v2) -> (price, iv, stat); raises RbError on failure."""
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    s, ps = _seed(seed)
    out = len(H) * len(T)
    price, iv, stat = np.zeros(out), np.zeros(out), np.zeros(out)
    cfg = _cfg(precision, compensator, method, scheme)
    check(_lib.load().rb_price_v2(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N), int(M), ps,
                                  len(s), ctypes.byref(cfg), price.ctypes.data_as(dp),
                                  iv.ctypes.data_as(dp), stat.ctypes.data_as(dp), out))
    return price, iv, stat


def iv_reference(price, K, T):
    """rb_iv_reference: implied vols to the reference's Boost tolerance (bracket by the factor 2
    from 0.2, midpoint of a bracket of relative width 2^-10)."""
    price, pp = _arr(price)
    K, pK = _arr(np.broadcast_to(np.asarray(K, dtype=np.float64), price.shape))
    T, pT = _arr(np.broadcast_to(np.asarray(T, dtype=np.float64), price.shape))
    iv = np.zeros(len(price))
    check(_lib.load().rb_iv_reference(pp, pK, pT, len(price), iv.ctypes.data_as(dp)))
    return iv


def price_noalloc(xi, H, eta, rho, T, K, N, M, seed, precision=_lib.PREC_SPLIT_FP16,
                  compensator=_lib.COMP_REF, method=_lib.METHOD_RT, stream=None):
    """rb_price_noalloc: rb_price_v2 on the fused kernel with no device allocation (a static
    per-device scratch array) -> (price, iv, stat); raises RbError on failure."""
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    s, ps = _seed(seed)
    out = len(H) * len(T)
    price, iv, stat = np.zeros(out), np.zeros(out), np.zeros(out)
    cfg = _cfg(precision, compensator, method, 0)
    check(_lib.load().rb_price_noalloc(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N), int(M), ps,
                                       len(s), ctypes.byref(cfg), stream, price.ctypes.data_as(dp),
                                       iv.ctypes.data_as(dp), stat.ctypes.data_as(dp), out))
    return price, iv, stat


def payoff_samples_v2(xi, H, eta, rho, T, K, N, Z, method=_lib.METHOD_RT,
                      compensator=_lib.COMP_REF, scheme=0):
    """Per-path payoffs from user Gaussians -> (payoff [M, out], price, stat)"""
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    Z, pZ = _arr(np.asarray(Z, dtype=np.float64).ravel())
    nst = 2 if method == _lib.METHOD_RT else 3
    M = Z.size // (nst * N)
    out = len(H) * len(T)
    payoff = np.zeros((M, out))
    price, stat = np.zeros(out), np.zeros(out)
    cfg = _cfg(_lib.PREC_FP64, compensator, method, scheme)
    check(_lib.load().rb_payoff_samples_v2(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N),
                                           int(M), pZ, ctypes.byref(cfg), payoff.ctypes.data_as(dp),
                                           price.ctypes.data_as(dp), stat.ctypes.data_as(dp)))
    return payoff, price, stat


def price_sobol(xi, H, eta, rho, T, K, N, M, precision=_lib.PREC_FP64, compensator=_lib.COMP_REF,
                method=_lib.METHOD_RT, scheme=0, seed=None, bridge=False):
    """Quasi-Monte Carlo pricing on Sobol' points (rb_price_sobol) -> (price, iv, stat).  seed:
    None for the reference's unscrambled points 1..M, else a random digital shift keyed by it.
    bridge: map each stream's N = 2^m dimensions through the Brownian-bridge construction
    (rb_price_sobol_bridge, hierarchical2increments of rBergomi/aux.h:105-125)."""
    s, ps = _seed(seed if seed is not None else [])
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    out = len(H) * len(T)
    price, iv, stat = np.zeros(out), np.zeros(out), np.zeros(out)
    cfg = _cfg(precision, compensator, method, scheme)
    fn = _lib.load().rb_price_sobol_bridge if bridge else _lib.load().rb_price_sobol
    check(fn(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N), int(M), ps, len(s), ctypes.byref(cfg),
             price.ctypes.data_as(dp), iv.ctypes.data_as(dp), stat.ctypes.data_as(dp), out))
    return price, iv, stat


def bridge_increments(Z):
    """The device Brownian-bridge mapping (rb_probe_bridge) of rows of N = 2^m values."""
    Z = np.ascontiguousarray(np.atleast_2d(np.asarray(Z, dtype=np.float64)))
    dW = np.zeros_like(Z)
    check(_lib.load().rb_probe_bridge(Z.ctypes.data_as(dp), Z.shape[0], Z.shape[1], dW.ctypes.data_as(dp)))
    return dW


def sobol_points(dims, i0, n):
    """Host Sobol' points i0 .. i0+n-1 of the quasi-Monte Carlo source: u[n][dims]."""
    u = np.zeros((n, dims))
    check(_lib.load().rb_probe_sobol(int(dims), int(i0), int(n), u.ctypes.data_as(dp)))
    return u


def price_rows(xi, H, eta, rho, T, K, N, M, seed, precision=_lib.PREC_FP64,
               compensator=_lib.COMP_REF, method=_lib.METHOD_RT, scheme=0):
    """Unordered rows (rb_price_rows): row i prices (T[i], K[i]) under (xi[i], H[i], eta[i],
    rho[i]) on shared paths; xi may be a scalar.  -> (price, iv, stat), one entry per row."""
    n = len(H)
    xi = np.full(n, float(xi)) if np.ndim(xi) == 0 else xi
    arrs = [_arr(x) for x in (xi, H, eta, rho, T, K)]
    if any(len(a) != n for a, _ in arrs):
        raise ValueError("every row array needs one entry per row")
    s, ps = _seed(seed)
    price, iv, stat = np.zeros(n), np.zeros(n), np.zeros(n)
    cfg = _cfg(precision, compensator, method, scheme)
    check(_lib.load().rb_price_rows(*[p for _, p in arrs], n, int(N), int(M), ps, len(s),
                                    ctypes.byref(cfg), price.ctypes.data_as(dp),
                                    iv.ctypes.data_as(dp), stat.ctypes.data_as(dp)))
    return price, iv, stat


def price_rows_turbo(xi, H, eta, rho, T, K, N, M, seed, m_control=10000, q_max=10000.0):
    """Turbo-charged RT estimator of unordered rows (rb_price_rows_turbo): control variate plus
    antithetic pairs, FP64 -> dict(price, iv, stat, alpha, Q, exceptions)."""
    n = len(H)
    xi = np.full(n, float(xi)) if np.ndim(xi) == 0 else xi
    arrs = [_arr(x) for x in (xi, H, eta, rho, T, K)]
    if any(len(a) != n for a, _ in arrs):
        raise ValueError("every row array needs one entry per row")
    s, ps = _seed(seed)
    out = {k: np.zeros(n) for k in ("price", "iv", "stat", "alpha", "Q")}
    exc = np.zeros(n, dtype=np.int64)
    P = lambda k: out[k].ctypes.data_as(dp)
    check(_lib.load().rb_price_rows_turbo(*[p for _, p in arrs], n, int(N), int(M), ps, len(s),
                                          int(m_control), float(q_max), P("price"), P("iv"),
                                          P("stat"), P("alpha"), P("Q"),
                                          exc.ctypes.data_as(ctypes.POINTER(ctypes.c_long))))
    out["exceptions"] = exc
    return out


def payoff_rows_samples(xi, H, eta, rho, T, K, N, Z, compensator=_lib.COMP_REF,
                        method=_lib.METHOD_RT, scheme=0):
    """Per-path payoffs of unordered rows from caller Gaussians Z[M][streams][N] (FP64 path)
    -> (payoff[M][rows], price, stat)."""
    n = len(H)
    xi = np.full(n, float(xi)) if np.ndim(xi) == 0 else xi
    arrs = [_arr(x) for x in (xi, H, eta, rho, T, K)]
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    M = Z.shape[0]
    payoff = np.zeros((M, n))
    price, stat = np.zeros(n), np.zeros(n)
    cfg = _cfg(_lib.PREC_FP64, compensator, method, scheme)
    check(_lib.load().rb_payoff_rows_samples(*[p for _, p in arrs], n, int(N), int(M),
                                             Z.ctypes.data_as(dp), ctypes.byref(cfg),
                                             payoff.ctypes.data_as(dp), price.ctypes.data_as(dp),
                                             stat.ctypes.data_as(dp)))
    return payoff, price, stat


class Plan:
    """Device-resident pricing plan (rb_plan_*): build once, run many times.

    stream: a HIP stream handle (int) to run on, e.g. torch.cuda.Stream().cuda_stream when torch
    was imported before this library (both then use one HIP runtime); None = the plan's own.
    graph: record the launches into a hipGraph on the first run and replay it afterwards.
    """

    def __init__(self, xi, H, eta, rho, T, K, N, M, seed, precision=_lib.PREC_FP64,
                 compensator=_lib.COMP_REF, method=_lib.METHOD_RT, profile=False, graph=False,
                 stream=None, scheme=0):
        self._keep = [_arr(H), _arr(eta), _arr(rho), _arr(T), _arr(K), _seed(seed)]
        (H, pH), (eta, pe), (rho, pr), (T, pT), (K, pK), (s, ps) = self._keep
        self.out_size = len(H) * len(T)
        self.cfg = _cfg(precision, compensator, method, scheme)
        self.ptr = ctypes.c_void_p()
        L = _lib.load()
        check(L.rb_plan_create(float(xi), pH, pe, pr, len(H), pT, pK, len(T), int(N), int(M), ps,
                               len(s), ctypes.byref(self.cfg), stream, ctypes.byref(self.ptr)))
        if profile:
            check(L.rb_plan_set_profiling(self.ptr, 1))
        if graph:
            check(L.rb_plan_set_graph(self.ptr, 1))

    def run(self):
        check(_lib.load().rb_plan_run(self.ptr))

    def results(self):
        n = self.out_size
        price, iv, stat = np.zeros(n), np.zeros(n), np.zeros(n)
        check(_lib.load().rb_plan_results(self.ptr, price.ctypes.data_as(dp), iv.ctypes.data_as(dp),
                                          stat.ctypes.data_as(dp)))
        return price, iv, stat

    def kernel_times(self):
        """{kernel name: (total ms, launches)} of the last run (profiling enabled)"""
        names = ctypes.create_string_buffer(32 * 16)
        ms = (ctypes.c_double * 16)()
        ln = (ctypes.c_int * 16)()
        cnt = ctypes.c_int()
        check(_lib.load().rb_plan_kernel_times(self.ptr, names, ms, ln, 16, ctypes.byref(cnt)))
        raw = names.raw
        return {raw[32 * i:32 * i + 32].split(b"\0")[0].decode(): (ms[i], ln[i])
                for i in range(cnt.value)}

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ch = ctypes.c_long()
        check(_lib.load().rb_plan_info(self.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                       ctypes.byref(ch)))
        return {"units": a.value, "groups": b.value, "options": c.value, "chunk": ch.value}

    def close(self):
        if self.ptr:
            _lib.load().rb_plan_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def layout(H, eta, rho, T, K, N, method=_lib.METHOD_RT):
    """Host-only planning: (nunits, ngroups, group_of_option, unit_of_group)"""
    H, pH = _arr(H)
    eta, pe = _arr(eta)
    rho, pr = _arr(rho)
    T, pT = _arr(T)
    K, pK = _arr(K)
    out = len(H) * len(T)
    nu, ng = ctypes.c_int(), ctypes.c_int()
    gofo = np.zeros(out, dtype=np.int32)
    uofg = np.zeros(out, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    check(_lib.load().rb_layout(pH, pe, pr, len(H), pT, pK, len(T), int(N), int(method),
                                ctypes.byref(nu), ctypes.byref(ng), gofo.ctypes.data_as(ip),
                                uofg.ctypes.data_as(ip)))
    return nu.value, ng.value, gofo, uofg[:ng.value]
