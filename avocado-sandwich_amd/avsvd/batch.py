"""Batched svdvals over the C ABI (avs_batch_*).

Many independent square matrices of mixed size go through one grouped
persistent band-reduction launch (every workgroup reduces whole matrices,
largest first, with the reference's Blockbidiag! schedule), then the host
stage (block bulge chasing + dbdsdc, FP64) runs on a thread pool, one matrix
per thread.  Sizes that are not multiples of the tile size are zero-padded;
the padding's zero singular values are dropped.
"""
import ctypes

import numpy as np

from ._lib import Opts, check, dptr, dtype_code, lib, vptr


class Batch:
    """Matrices of orders `ns` reduced together; b=None takes the tile size
    from the autotuner's table (form "batch", keyed by the largest order).
    update_mode: UPD_SPLIT_BF16 runs FP32 trailing updates as split-BF16."""

    def __init__(self, ns, dtype=np.float32, b=None, workers=0, update_mode=0):
        self.ns = [int(x) for x in ns]
        self.dtype = np.dtype(dtype)
        if b is None:
            from .autotune import choose_tile
            b = choose_tile(max(self.ns), self.dtype, "batch")[0]
        self.b = b
        self._h = ctypes.c_void_p()
        o = Opts()
        lib().avs_default_opts(ctypes.byref(o))
        o.workers = workers
        o.update_mode = update_mode
        arr = (ctypes.c_int64 * len(self.ns))(*self.ns)
        check(lib().avs_batch_create(ctypes.byref(self._h), dtype_code(self.dtype), len(self.ns), arr, b,
                                     ctypes.byref(o)))
        self._keep = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().avs_batch_destroy(h)
            self._h = None

    def set(self, idx, A):
        A = np.asfortranarray(A, dtype=self.dtype)
        assert A.shape == (self.ns[idx], self.ns[idx])
        self._keep[idx] = A  # the upload is asynchronous
        check(lib().avs_batch_set_matrix_host(self._h, idx, vptr(A), A.shape[0]))

    def blockbidiag(self):
        check(lib().avs_batch_blockbidiag(self._h))

    def last_time_ms(self):
        ms = ctypes.c_double()
        check(lib().avs_batch_last_time_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def band(self, idx):
        N = -(-self.ns[idx] // self.b) * self.b
        out = np.zeros((self.b + 1, N), order="F")
        check(lib().avs_batch_get_band(self._h, idx, dptr(out), self.b + 1))
        return out

    def svdvals(self):
        out = np.empty(sum(self.ns))
        check(lib().avs_batch_svdvals(self._h, dptr(out)))
        self._keep.clear()
        res, o = [], 0
        for n in self.ns:
            res.append(out[o:o + n])
            o += n
        return res


def batch_svdvals(mats, dtype=np.float32, b=None):
    """Singular values of every matrix in `mats` (list of square arrays)."""
    bt = Batch([m.shape[0] for m in mats], dtype=dtype, b=b)
    for i, m in enumerate(mats):
        bt.set(i, m)
    return bt.svdvals()
