"""ctypes binding of libavsvd.so (the C ABI in include/avsvd.h).

The shared library is built in-tree (avocado-sandwich_amd/libavsvd.so) by
__graft_entry__.build() / `make -C avocado-sandwich_amd/csrc`.  Loading fails
loudly if it is missing: there is no CPU fallback for the device path.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the in-tree build, unless AVSVD_LIBRARY names another build of the same library
LIB_PATH = os.environ.get("AVSVD_LIBRARY") or os.path.join(PKG_DIR, "libavsvd.so")

AVS_F64, AVS_F32 = 0, 1
SCHED_PERSISTENT, SCHED_HOST, SCHED_STREAMS = 0, 1, 2
TREE_FLAT, TREE_HIER = 0, 1
UPD_NATIVE, UPD_SPLIT_BF16 = 0, 1

# every symbol include/avsvd.h declares
EXPORTS = (
    "avs_default_opts", "avs_create", "avs_destroy", "avs_set_matrix_host", "avs_set_matrix_device",
    "avs_get_matrix_host", "avs_get_matrix_device", "avs_blockbidiag", "avs_tiled_qr", "avs_get_band",
    "avs_svdvals", "avs_synchronize", "avs_last_time_ms", "avs_last_error", "avs_bidiagonalize_band",
    "avs_bidiagonalize_band_lapack", "avs_bidiagonalize_band_chase", "avs_bidiagonalize_band_device",
    "avs_bdsdc_values", "avs_geqrt", "avs_tsqrt", "avs_unmqr", "avs_tsmqr", "avs_tsmqr_noalloc", "avs_gen_gaussian_host",
    "avs_debug_mfma_layout", "avs_schedule_stats", "avs_schedule_dump", "avs_phase_dump", "avs_phase_timing", "avs_set_phase_timing",
    "avs_phase_trace",
    "avs_batch_create", "avs_batch_destroy", "avs_batch_set_matrix_host", "avs_batch_blockbidiag",
    "avs_batch_last_time_ms", "avs_batch_get_band", "avs_batch_svdvals", "avs_set_matrix_host_sub",
    "avs_server_default_opts", "avs_server_create", "avs_server_submit", "avs_server_poll", "avs_server_wait",
    "avs_server_stats", "avs_server_destroy", "avs_tune_set", "avs_tune_clear", "avs_tune_lookup",
    "avs_device_arch", "avs_ooc_blockbidiag", "avs_ooc_tiled_qr", "avs_apply_q",
)


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
                ("host_threads", ctypes.c_int32), ("b_f32", ctypes.c_int32), ("b_f64", ctypes.c_int32),
                ("update_mode", ctypes.c_int32), ("device_sigma", ctypes.c_int32), ("reserved", ctypes.c_int32 * 6)]


class Opts(ctypes.Structure):
    _fields_ = [("stream", ctypes.c_void_p), ("sched", ctypes.c_int32), ("tree", ctypes.c_int32),
                ("domain", ctypes.c_int32), ("update_mode", ctypes.c_int32), ("workers", ctypes.c_int32),
                ("wait_ms", ctypes.c_int32), ("phase_timing", ctypes.c_int32), ("single_stream", ctypes.c_int32),
                ("reserve_cus", ctypes.c_int32), ("graph", ctypes.c_int32), ("masked_modes", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 3)]


class AvsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("avsvd status %d: %s" % (status, msg))
        self.status = status


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libavsvd.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        # torch ships its own HIP runtime; once another one is loaded first,
        # torch's device initialisation breaks the process's HIP state.  Load
        # torch (when installed) before the library so both use torch's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
        L.avs_last_error.restype = ctypes.c_char_p
        L.avs_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.avs_default_opts.restype = None
        L.avs_server_default_opts.argtypes = [ctypes.POINTER(ServerOpts)]
        L.avs_server_default_opts.restype = None
        sig = {
            "avs_create": [ctypes.POINTER(vp), i32, i64, i32, ctypes.POINTER(Opts)],
            "avs_destroy": [vp],
            "avs_set_matrix_host": [vp, vp, i64],
            "avs_set_matrix_device": [vp, vp, i64],
            "avs_get_matrix_host": [vp, vp, i64],
            "avs_get_matrix_device": [vp, vp, i64],
            "avs_blockbidiag": [vp, i32],
            "avs_tiled_qr": [vp, i32],
            "avs_get_band": [vp, dp, i64],
            "avs_svdvals": [vp, dp],
            "avs_synchronize": [vp],
            "avs_last_time_ms": [vp, dp],
            "avs_bidiagonalize_band": [i64, i32, dp, i64, dp, dp],
            "avs_bidiagonalize_band_lapack": [i64, i32, dp, i64, dp, dp],
            "avs_bidiagonalize_band_chase": [i64, i32, dp, i64, dp, dp],
            "avs_bidiagonalize_band_device": [i64, i32, dp, i64, dp, dp],
            "avs_bdsdc_values": [i64, dp, dp, dp],
            "avs_geqrt": [i32, i32, vp, i64, vp, i32, vp],
            "avs_tsqrt": [i32, i32, vp, i64, vp, i64, vp, i32, vp],
            "avs_unmqr": [i32, i32, vp, i64, vp, vp, i64, i64, i32, i32, i32, vp],
            "avs_tsmqr": [i32, i32, vp, i64, vp, i64, vp, i64, vp, i64, i32, i32, i32, vp],
            "avs_tsmqr_noalloc": [i32, i32, vp, i64, vp, i64, vp, i64, vp, i64, i32, i32, i32, vp],
            "avs_ooc_blockbidiag": [i32, i64, i32, vp, i64, i32, ctypes.POINTER(Opts)],
            "avs_ooc_tiled_qr": [i32, i64, i32, vp, i64, i32, ctypes.POINTER(Opts)],
            "avs_apply_q": [vp, vp, i64, i64, i32],
            "avs_gen_gaussian_host": [ctypes.c_uint64, i64, ctypes.c_double, i32, vp, i64],
            "avs_debug_mfma_layout": [i32, dp, ctypes.POINTER(ctypes.c_float)],
            "avs_schedule_stats": [i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(i64), dp],
            "avs_schedule_dump": [i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_int32), i64,
                                  ctypes.POINTER(i64)],
            "avs_phase_timing": [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64),
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_double)],
            "avs_phase_dump": [i32, i32, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_int32), i64,
                               ctypes.POINTER(i64)],
            "avs_set_phase_timing": [vp, i32],
            "avs_phase_trace": [vp, dp, dp, ctypes.POINTER(ctypes.c_int32), i64, ctypes.POINTER(i64)],
            "avs_batch_create": [ctypes.POINTER(vp), i32, i32, ctypes.POINTER(i64), i32, ctypes.POINTER(Opts)],
            "avs_batch_destroy": [vp],
            "avs_batch_set_matrix_host": [vp, i32, vp, i64],
            "avs_batch_blockbidiag": [vp],
            "avs_batch_last_time_ms": [vp, dp],
            "avs_batch_get_band": [vp, i32, dp, i64],
            "avs_batch_svdvals": [vp, dp],
            "avs_set_matrix_host_sub": [vp, vp, i64, i64],
            "avs_server_create": [ctypes.POINTER(vp), ctypes.POINTER(ServerOpts)],
            "avs_server_submit": [vp, i32, i64, vp, i64, ctypes.POINTER(ctypes.c_uint64)],
            "avs_server_poll": [vp, ctypes.c_uint64, ctypes.POINTER(i32)],
            "avs_server_wait": [vp, ctypes.c_uint64, dp, dp],
            "avs_server_stats": [vp, dp, dp, dp, ctypes.POINTER(i64)],
            "avs_server_destroy": [vp],
            "avs_tune_set": [i32, i64, i32],
            "avs_tune_clear": [],
            "avs_tune_lookup": [i32, i64, ctypes.POINTER(i32)],
            "avs_device_arch": [i32, ctypes.c_char_p, i64],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int32
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        raise AvsError(status, lib().avs_last_error().decode(errors="replace"))


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def vptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def dtype_code(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        return AVS_F64
    if dt == np.float32:
        return AVS_F32
    raise TypeError("unsupported element type %s (Float64 and Float32 are built)" % dt)
