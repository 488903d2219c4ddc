"""Continuous-batching svdvals request loop (avs_server_* in the C ABI).

Requests of mixed n and precision are submitted as they arrive; the library
admits them to the device against an HBM budget, overlaps their band
reductions on separate streams, ships only the band to the host and runs the
host stage on a thread pool.  Latency is measured from submit to the
singular values being ready.
"""
import ctypes

import numpy as np

from ._lib import ServerOpts, check, dptr, dtype_code, lib, vptr


def served_tile(dtype, n):
    """Tile size the serving loop gives an order-n request of `dtype` whose
    tile size is not fixed (the library's tile-size table)."""
    b = ctypes.c_int32()
    check(lib().avs_tune_lookup(dtype_code(np.dtype(dtype)), int(n), ctypes.byref(b)))
    return b.value


class Server:
    """Serving loop.  b_f32 / b_f64 fix the tile size per dtype; 0 leaves it to
    the autotuner's table (`tune_table`: a path, default the table
    avsvd.autotune.load_table() finds), which is loaded into the library here.
    device_sigma: the bidiagonals' singular values by bisection on the device
    (per host worker, buffers allocated once) instead of LAPACK dbdsdc."""

    def __init__(self, hbm_budget_gb=64, max_inflight=4, host_threads=0, b_f32=0, b_f64=0, update_mode=0,
                 tune_table=None, device_sigma=False):
        if not (b_f32 and b_f64):
            from .autotune import load_table, push_to_library
            push_to_library(load_table(tune_table))
        o = ServerOpts()
        lib().avs_server_default_opts(ctypes.byref(o))
        o.hbm_budget_bytes = int(hbm_budget_gb * (1 << 30))
        o.max_inflight = max_inflight
        if host_threads:
            o.host_threads = host_threads
        o.b_f32, o.b_f64, o.update_mode = b_f32, b_f64, update_mode
        o.device_sigma = 1 if device_sigma else 0
        self._h = ctypes.c_void_p()
        check(lib().avs_server_create(ctypes.byref(self._h), ctypes.byref(o)))
        self._n = {}

    def close(self):
        if self._h is not None and self._h.value:
            lib().avs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, A):
        A = np.asfortranarray(A)
        rid = ctypes.c_uint64()
        check(lib().avs_server_submit(self._h, dtype_code(A.dtype), A.shape[0], vptr(A), A.shape[0],
                                      ctypes.byref(rid)))
        self._n[rid.value] = A.shape[0]
        return rid.value

    def done(self, rid):
        d = ctypes.c_int32()
        check(lib().avs_server_poll(self._h, rid, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, rid):
        """(sigma, latency_ms) of request rid."""
        out = np.empty(self._n.pop(rid))
        lat = ctypes.c_double()
        check(lib().avs_server_wait(self._h, rid, dptr(out), ctypes.byref(lat)))
        return out, lat.value

    def stats(self):
        p50, p99, sps = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().avs_server_stats(self._h, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(sps),
                                     ctypes.byref(n)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "solves_per_s": sps.value, "completed": n.value}
