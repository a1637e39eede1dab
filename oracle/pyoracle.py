"""ctypes wrapper of the CPU oracle (oracle/liborbergomi.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker.  The product package rbergomi_amd never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborbergomi.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)

COMP_REF, COMP_EXACT = 0, 1
CONV_DIRECT, CONV_FFT = 0, 1

_lib = None


def build(force=False):
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, l = ctypes.c_double, ctypes.c_int, ctypes.c_long
        L.orc_gamma.argtypes = [d, i, _dp]
        L.orc_convolve_direct.argtypes = [_dp, _dp, i, _dp]
        L.orc_convolve_fft.argtypes = [_dp, _dp, i, _dp]
        L.orc_wtilde.argtypes = [_dp, _dp, d, i, i, _dp]
        L.orc_update_v.argtypes = [_dp, d, d, d, d, i, i, _dp]
        L.orc_int_vdt.argtypes = [_dp, i, d]
        L.orc_int_vdt.restype = d
        L.orc_int_rootvdw.argtypes = [_dp, _dp, i, d]
        L.orc_int_rootvdw.restype = d
        for fn in ("orc_phi",):
            getattr(L, fn).argtypes = [d]
            getattr(L, fn).restype = d
        for fn in ("orc_bs_call", "orc_bs_put", "orc_iv_call", "orc_iv_call_reference"):
            getattr(L, fn).argtypes = [d, d, d, d]
            getattr(L, fn).restype = d
        L.orc_rt_payoff.argtypes = [d, d, d, d]
        L.orc_rt_payoff.restype = d
        samp = [d, _dp, _dp, _dp, i, _dp, _dp, i, i, l, _dp, i, _dp]
        L.orc_payoff_rt_samples.argtypes = samp
        L.orc_integrals_rt_samples.argtypes = samp
        L.orc_payoff_mc_samples.argtypes = samp
        L.orc_payoff_rows_samples.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i, i, l, _dp, i, _dp]
        L.orc_hyp2f1.argtypes = [d, d, d, d]
        L.orc_hyp2f1.restype = d
        L.orc_chol_cov.argtypes = [d, i, _dp]
        L.orc_chol_factor.argtypes = [d, i, _dp]
        L.orc_payoff_chol_samples.argtypes = [d, _dp, _dp, _dp, i, _dp, _dp, i, i, l, _dp, i, i, _dp]
        L.orc_turbo_rows.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i, i, l, _dp, l, _dp, d, i,
                                     _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_long)]
        L.orc_price_stat.argtypes = [_dp, l, i, _dp, _dp]
        seeded = [d, _dp, _dp, _dp, i, _dp, _dp, i, i, l, _u64p, i, i, i, _dp, _dp, _dp]
        L.orc_compute_rt.argtypes = seeded
        L.orc_compute_mc.argtypes = seeded
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_philox_key.argtypes = [_u64p, i, _u32p]
        L.orc_normals_philox.argtypes = [_u64p, i, l, l, i, i, _dp]
        L.orc_bridge.argtypes = [_dp, i, _dp]
        L.orc_bridge.restype = None
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def gamma(H, N):
    out = np.zeros(N)
    lib().orc_gamma(H, N, out.ctypes.data_as(_dp))
    return out


def convolve(x, y, fft=False):
    x, px = _d(x)
    y, py = _d(y)
    z = np.zeros(len(x))
    (lib().orc_convolve_fft if fft else lib().orc_convolve_direct)(px, py, len(x), z.ctypes.data_as(_dp))
    return z


def wtilde(W1, W1perp, H, conv=CONV_DIRECT):
    W1, p1 = _d(W1)
    W1perp, p2 = _d(W1perp)
    out = np.zeros(len(W1))
    lib().orc_wtilde(p1, p2, H, len(W1), conv, out.ctypes.data_as(_dp))
    return out


def update_v(Wt, xi, H, eta, T, comp=COMP_REF):
    Wt, p = _d(Wt)
    out = np.zeros(len(Wt))
    lib().orc_update_v(p, xi, H, eta, T, len(Wt), comp, out.ctypes.data_as(_dp))
    return out


def phi(x):
    return lib().orc_phi(float(x))


def bs_call(S0, K, tau, sigma):
    return lib().orc_bs_call(S0, K, tau, sigma)


def bs_put(S0, K, tau, sigma):
    return lib().orc_bs_put(S0, K, tau, sigma)


def iv_call(price, S0, K, tau):
    return lib().orc_iv_call(price, S0, K, tau)


def iv_call_reference(price, S0, K, tau):
    return lib().orc_iv_call_reference(price, S0, K, tau)


def rt_payoff(Iv, Is, rho, K):
    return lib().orc_rt_payoff(Iv, Is, rho, K)


def _params(H, eta, rho, T, K):
    H, pH = _d(np.atleast_1d(H))
    eta, pe = _d(np.atleast_1d(eta))
    rho, pr = _d(np.atleast_1d(rho))
    T, pT = _d(np.atleast_1d(T))
    K, pK = _d(np.atleast_1d(K))
    assert len(H) == len(eta) == len(rho) and len(T) == len(K)
    return (H, eta, rho, T, K), (pH, pe, pr, len(H), pT, pK, len(T))


def payoff_rt_samples(xi, H, eta, rho, T, K, N, Z, comp=COMP_REF):
    """Z: [M, 2, N] (or flat 2*N*M) -> payoff [M, out_size]"""
    keep, args = _params(H, eta, rho, T, K)
    Z, pZ = _d(Z)
    M = Z.size // (2 * N)
    out = np.zeros((M, args[3] * args[6]))
    lib().orc_payoff_rt_samples(xi, *args, N, M, pZ, comp, out.ctypes.data_as(_dp))
    return out


def integrals_rt_samples(xi, H, eta, rho, T, K, N, Z, comp=COMP_REF):
    """-> [M, out_size, 2] with (Ivdt, IsvdW)"""
    keep, args = _params(H, eta, rho, T, K)
    Z, pZ = _d(Z)
    M = Z.size // (2 * N)
    out = np.zeros((M, args[3] * args[6], 2))
    lib().orc_integrals_rt_samples(xi, *args, N, M, pZ, comp, out.ctypes.data_as(_dp))
    return out


def payoff_mc_samples(xi, H, eta, rho, T, K, N, Z, comp=COMP_REF):
    """Z: [M, 3, N] -> payoff [M, out_size]"""
    keep, args = _params(H, eta, rho, T, K)
    Z, pZ = _d(Z)
    M = Z.size // (3 * N)
    out = np.zeros((M, args[3] * args[6]))
    lib().orc_payoff_mc_samples(xi, *args, N, M, pZ, comp, out.ctypes.data_as(_dp))
    return out


def payoff_rows_samples(xi, H, eta, rho, T, K, N, Z, comp=COMP_REF):
    """unordered rows: payoff[m][i] of row i (xi may be a scalar)"""
    n = len(H)
    xi = np.full(n, float(xi)) if np.ndim(xi) == 0 else xi
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (xi, H, eta, rho, T, K)]
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    M = Z.shape[0]
    out = np.zeros((M, n))
    lib().orc_payoff_rows_samples(*[a.ctypes.data_as(_dp) for a in arrs], n, N, M,
                                  Z.ctypes.data_as(_dp), comp, out.ctypes.data_as(_dp))
    return out


def turbo_rows(xi, H, eta, rho, T, K, N, Zc, Z, qmax=10000.0, comp=COMP_REF):
    """turbo-charged RT estimator of unordered rows from pilot Gaussians Zc[Mc][2][N] and main
    Gaussians Z[M][2][N] -> dict(price, stat, alpha, Q, exceptions)"""
    n = len(H)
    xi = np.full(n, float(xi)) if np.ndim(xi) == 0 else xi
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (xi, H, eta, rho, T, K)]
    Zc = np.ascontiguousarray(Zc, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("price", "stat", "alpha", "Q")}
    exc = np.zeros(n, dtype=np.int64)
    lib().orc_turbo_rows(*[a.ctypes.data_as(_dp) for a in arrs], n, N, Zc.shape[0],
                         Zc.ctypes.data_as(_dp), Z.shape[0], Z.ctypes.data_as(_dp), float(qmax), comp,
                         *[out[k].ctypes.data_as(_dp) for k in ("price", "stat", "alpha", "Q")],
                         exc.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
    out["exceptions"] = exc
    return out


def hyp2f1(a, b, c, z):
    return lib().orc_hyp2f1(a, b, c, z)


def chol_cov(H, n):
    A = np.zeros((2 * n, 2 * n))
    lib().orc_chol_cov(H, n, A.ctypes.data_as(_dp))
    return A


def chol_factor(H, n):
    L = np.zeros((2 * n, 2 * n))
    if lib().orc_chol_factor(H, n, L.ctypes.data_as(_dp)) != 0:
        raise ValueError("covariance not positive definite")
    return L


def payoff_chol_samples(xi, H, eta, rho, T, K, N, Z, comp=COMP_REF, method=0):
    """per-path payoffs with the exact (Cholesky) Wtilde; Z[M][2 or 3][N]"""
    keep, args = _params(H, eta, rho, T, K)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    M = Z.shape[0]
    pay = np.zeros((M, args[3] * args[6]))
    rc = lib().orc_payoff_chol_samples(xi, *args, N, M, Z.ctypes.data_as(_dp), comp, method,
                                       pay.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("covariance not positive definite")
    return pay


def price_stat(payoff):
    payoff, p = _d(payoff)
    M, out = payoff.shape
    price, stat = np.zeros(out), np.zeros(out)
    lib().orc_price_stat(p, M, out, price.ctypes.data_as(_dp), stat.ctypes.data_as(_dp))
    return price, stat


def _seeded(fn, xi, H, eta, rho, T, K, N, M, seed, num_threads, conv):
    keep, args = _params(H, eta, rho, T, K)
    seed = np.ascontiguousarray(np.atleast_1d(seed), dtype=np.uint64)
    out = args[3] * args[6]
    price, iv, stat = np.zeros(out), np.zeros(out), np.zeros(out)
    fn(xi, *args, N, M, seed.ctypes.data_as(_u64p), len(seed), num_threads, conv,
       price.ctypes.data_as(_dp), iv.ctypes.data_as(_dp), stat.ctypes.data_as(_dp))
    return price, iv, stat


def compute_rt(xi, H, eta, rho, T, K, N, M, seed, num_threads=1, conv=CONV_FFT):
    return _seeded(lib().orc_compute_rt, xi, H, eta, rho, T, K, N, M, seed, num_threads, conv)


def compute_mc(xi, H, eta, rho, T, K, N, M, seed, num_threads=1, conv=CONV_FFT):
    return _seeded(lib().orc_compute_mc, xi, H, eta, rho, T, K, N, M, seed, num_threads, conv)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data_as(_u32p), k.ctypes.data_as(_u32p), out.ctypes.data_as(_u32p))
    return out


def philox_key(seed):
    s = np.ascontiguousarray(np.atleast_1d(seed), dtype=np.uint64)
    out = np.zeros(2, dtype=np.uint32)
    lib().orc_philox_key(s.ctypes.data_as(_u64p), len(s), out.ctypes.data_as(_u32p))
    return out


def bridge(Z):
    """hierarchical2increments of each row of Z (rows of 2^m values) -> dW, same shape"""
    Z = np.ascontiguousarray(np.atleast_2d(np.asarray(Z, dtype=np.float64)))
    n = Z.shape[1]
    m = int(n).bit_length() - 1
    assert 1 << m == n, "rows of 2^m values"
    dW = np.zeros_like(Z)
    for r in range(Z.shape[0]):
        _, pz = _d(Z[r])
        lib().orc_bridge(pz, m, dW[r].ctypes.data_as(_dp))
    return dW


def normals_philox(seed, p0, np_, nstreams, N):
    """-> [np_, nstreams, N] standard normals of paths p0..p0+np_-1"""
    s = np.ascontiguousarray(np.atleast_1d(seed), dtype=np.uint64)
    out = np.zeros((np_, nstreams, N))
    lib().orc_normals_philox(s.ctypes.data_as(_u64p), len(s), p0, np_, nstreams, N,
                             out.ctypes.data_as(_dp))
    return out
