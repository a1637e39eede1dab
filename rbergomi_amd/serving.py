"""Batched pricing and the continuous-batching server (rb_price_batch, rb_server_*).

Each request is the reference's tensor form (xi, H, eta, rho zipped; T, K zipped; N, M, seed),
as accepted by computeIVRT_tensor (rBergLib/cinterface.h:68-71), priced on the split-precision
kernel.  A request's results are bit-identical whether it is priced alone or in a batch.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import RbConfig, check, dp, u64p


class RbRequest(ctypes.Structure):
    _fields_ = [("xi", ctypes.c_double), ("H", dp), ("eta", dp), ("rho", dp), ("par_size", ctypes.c_int),
                ("T", dp), ("K", dp), ("op_size", ctypes.c_int), ("N", ctypes.c_int), ("M", ctypes.c_long),
                ("seed", u64p), ("seed_size", ctypes.c_int), ("price", dp), ("iv", dp), ("stat", dp)]


class RbServerConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("max_inflight", ctypes.c_int),
                ("max_batch_requests", ctypes.c_int), ("max_device_bytes", ctypes.c_long),
                ("max_batch_path_steps", ctypes.c_double)]


class RbServerReport(ctypes.Structure):
    _fields_ = [("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("completed", ctypes.c_long),
                ("path_steps", ctypes.c_double), ("batches", ctypes.c_long), ("peak_inflight", ctypes.c_long),
                ("kernel_launches", ctypes.c_long), ("kernel_ms", ctypes.c_double),
                ("device_busy_ms", ctypes.c_double), ("device_overlap_ms", ctypes.c_double),
                ("device_span_ms", ctypes.c_double), ("max_concurrent", ctypes.c_long),
                ("conv_flops", ctypes.c_double)]


def _declare(L):
    if getattr(L, "_serving_declared", False):
        return
    vp = ctypes.c_void_p
    L.rb_price_batch.argtypes = [ctypes.POINTER(RbRequest), ctypes.c_int, ctypes.POINTER(RbConfig)]
    L.rb_batch_create.argtypes = [ctypes.POINTER(RbRequest), ctypes.c_int, ctypes.POINTER(RbConfig), vp,
                                  ctypes.POINTER(vp)]
    L.rb_batch_update.argtypes = [vp, ctypes.POINTER(RbRequest), ctypes.c_int]
    L.rb_batch_set_graph.argtypes = [vp, ctypes.c_int]
    L.rb_batch_run.argtypes = [vp]
    L.rb_batch_destroy.argtypes = [vp]
    L.rb_batch_destroy.restype = None
    L.rb_server_create.argtypes = [ctypes.POINTER(RbServerConfig), ctypes.POINTER(vp)]
    L.rb_server_create_method.argtypes = [ctypes.POINTER(RbServerConfig), ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(vp)]
    L.rb_server_report_get.argtypes = [vp, ctypes.POINTER(RbServerReport)]
    L.rb_server_set_timing.argtypes = [vp, ctypes.c_int]
    L.rb_server_submit.argtypes = [vp, ctypes.POINTER(RbRequest), ctypes.POINTER(ctypes.c_long)]
    L.rb_server_wait.argtypes = [vp, ctypes.c_long, dp]
    L.rb_server_stats.argtypes = [vp, dp, dp, ctypes.POINTER(ctypes.c_long), dp, ctypes.POINTER(ctypes.c_long)]
    L.rb_server_destroy.argtypes = [vp]
    L.rb_server_destroy.restype = None
    ip = ctypes.POINTER(ctypes.c_int)
    L.rb_autotune_fused.argtypes = [ctypes.POINTER(RbRequest), ip, ctypes.c_int, ctypes.POINTER(RbConfig), ip, dp]
    L.rb_tune_save.argtypes = [ctypes.c_char_p]
    L.rb_tune_load.argtypes = [ctypes.c_char_p]
    L.rb_tune_clear.argtypes = []
    L.rb_tune_table.argtypes = [ip, ctypes.c_int]
    L._serving_declared = True


class Request:
    """One pricing request with its own output arrays (kept alive until done)."""

    def __init__(self, xi, H, eta, rho, T, K, N, M, seed):
        a = lambda x: np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
        self.H, self.eta, self.rho, self.T, self.K = map(a, (H, eta, rho, T, K))
        self.seed = np.ascontiguousarray(np.atleast_1d(np.asarray(seed, dtype=np.uint64)))
        n = len(self.H) * len(self.T)
        self.price, self.iv, self.stat = np.zeros(n), np.zeros(n), np.zeros(n)
        self.N, self.M, self.xi = int(N), int(M), float(xi)
        P = lambda x: x.ctypes.data_as(dp)
        self.c = RbRequest(self.xi, P(self.H), P(self.eta), P(self.rho), len(self.H), P(self.T), P(self.K),
                           len(self.T), self.N, self.M, self.seed.ctypes.data_as(u64p), len(self.seed),
                           P(self.price), P(self.iv), P(self.stat))
        self.ticket = None
        self.latency_ms = None


def price_batch(requests, precision=_lib.PREC_SPLIT_FP16, method=_lib.METHOD_RT,
                compensator=_lib.COMP_REF):
    """Prices all requests with one launch; fills each request's price/iv/stat."""
    L = _lib.load()
    _declare(L)
    arr = (RbRequest * len(requests))(*[r.c for r in requests])
    cfg = RbConfig(precision, compensator, method, 0)
    check(L.rb_price_batch(arr, len(requests), ctypes.byref(cfg)))
    return requests


class Batch:
    """A set of requests planned once and priced many times (rb_batch_*): run() fills every
    request's price/iv/stat; update(new_requests) swaps in new parameters of the same shapes;
    graph=True records the launches into a hipGraph and replays it."""

    def __init__(self, requests, precision=_lib.PREC_SPLIT_FP16, method=_lib.METHOD_RT,
                 compensator=_lib.COMP_REF, graph=False, stream=None):
        L = _lib.load()
        _declare(L)
        self.requests = list(requests)
        self._arr = (RbRequest * len(self.requests))(*[r.c for r in self.requests])
        self.cfg = RbConfig(precision, compensator, method, 0)
        self.ptr = ctypes.c_void_p()
        check(L.rb_batch_create(self._arr, len(self.requests), ctypes.byref(self.cfg), stream,
                                ctypes.byref(self.ptr)))
        if graph:
            check(L.rb_batch_set_graph(self.ptr, 1))

    def run(self):
        check(_lib.load().rb_batch_run(self.ptr))
        return self.requests

    def update(self, requests):
        requests = list(requests)
        arr = (RbRequest * len(requests))(*[r.c for r in requests])
        check(_lib.load().rb_batch_update(self.ptr, arr, len(requests)))
        self.requests, self._arr = requests, arr

    def close(self):
        if self.ptr:
            _lib.load().rb_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(shape, Ns, precision=_lib.PREC_SPLIT_FP16, method=_lib.METHOD_RT, compensator=_lib.COMP_REF):
    """Times the fused kernel's tile shapes for each N in Ns with the request shape `shape` (a
    Request; its N is ignored) and records the fastest for later plans, batches and server
    requests of this process (rb_autotune_fused).  Returns {N: (path blocks per tile, ms)}."""
    L = _lib.load()
    _declare(L)
    Ns = np.ascontiguousarray(np.asarray(Ns, dtype=np.int32))
    pb = np.zeros(len(Ns), dtype=np.int32)
    ms = np.zeros(len(Ns))
    ip = ctypes.POINTER(ctypes.c_int)
    cfg = RbConfig(precision, compensator, method, 0)
    check(L.rb_autotune_fused(ctypes.byref(shape.c), Ns.ctypes.data_as(ip), len(Ns), ctypes.byref(cfg),
                              pb.ctypes.data_as(ip), ms.ctypes.data_as(dp)))
    return {int(n): (int(p), float(t)) for n, p, t in zip(Ns, pb, ms)}


def tune_table():
    """The tuned entries as (padded N, operand mode, plain MC, path blocks per tile) tuples."""
    L = _lib.load()
    _declare(L)
    n = L.rb_tune_table(None, 0)
    rows = np.zeros(4 * max(n, 1), dtype=np.int32)
    L.rb_tune_table(rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    return [tuple(int(v) for v in rows[4 * i:4 * i + 4]) for i in range(n)]


def tune_save(path):
    L = _lib.load()
    _declare(L)
    check(L.rb_tune_save(str(path).encode()))


def tune_load(path):
    L = _lib.load()
    _declare(L)
    check(L.rb_tune_load(str(path).encode()))


def tune_clear():
    L = _lib.load()
    _declare(L)
    check(L.rb_tune_clear())


class Server:
    """Continuous-batching server (rb_server_*)."""

    def __init__(self, precision=_lib.PREC_SPLIT_FP16, max_inflight=2, max_batch_requests=64,
                 max_device_bytes=8 << 30, max_batch_path_steps=0.0, method=_lib.METHOD_RT,
                 compensator=_lib.COMP_REF):
        L = _lib.load()
        _declare(L)
        self.ptr = ctypes.c_void_p()
        cfg = RbServerConfig(precision, max_inflight, max_batch_requests, max_device_bytes,
                             float(max_batch_path_steps))
        check(L.rb_server_create_method(ctypes.byref(cfg), method, compensator, ctypes.byref(self.ptr)))
        self.live = {}

    def submit(self, req):
        t = ctypes.c_long()
        check(_lib.load().rb_server_submit(self.ptr, ctypes.byref(req.c), ctypes.byref(t)))
        req.ticket = t.value
        self.live[t.value] = req
        return req

    def wait(self, req):
        lat = ctypes.c_double()
        rc = _lib.load().rb_server_wait(self.ptr, req.ticket, ctypes.byref(lat))
        self.live.pop(req.ticket, None)
        check(rc)
        req.latency_ms = lat.value
        return req

    def stats(self):
        p50, p99, ps = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n, b = ctypes.c_long(), ctypes.c_long()
        check(_lib.load().rb_server_stats(self.ptr, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(n),
                                          ctypes.byref(ps), ctypes.byref(b)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "completed": n.value, "path_steps": ps.value,
                "batches": b.value}

    def set_timing(self, enable):
        """Timing events around the batches' kernels from now on (kernel intervals in report())."""
        check(_lib.load().rb_server_set_timing(self.ptr, 1 if enable else 0))

    def report(self):
        """rb_server_report_get as a dict: latency, work, and the device timeline of the batches'
        pricing kernels (union, overlap, most at once)."""
        r = RbServerReport()
        check(_lib.load().rb_server_report_get(self.ptr, ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in RbServerReport._fields_}

    def close(self):
        if self.ptr:
            for r in list(self.live.values()):
                try:
                    self.wait(r)
                except Exception:
                    pass
            _lib.load().rb_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
