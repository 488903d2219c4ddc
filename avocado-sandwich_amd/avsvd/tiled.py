"""TiledMatrix and the tiled algorithms over the C ABI.

Each function cites the reference function it stands in for (paths relative
to the reference repository root).  Device work runs in libavsvd.so; this
module only moves host arrays across the boundary and keeps the reference's
names, argument meaning and error behaviour.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AvsError, Opts, check, dptr, dtype_code, lib, vptr

_compiled = {"tilesize": 32, "tilefactor": 1, "tiledim2d": False}
_BUILT_TILES = (32, 64, 128)


def compile_kernels(tilesize=32, tilefactor=1, tiledim2d=False):
    """compile_kernels (tiledalgos.jl:262-267).  The gfx950 kernels are compiled
    ahead of time for b in {32, 64, 128}; this records the default tile size
    and checks it is built (the reference re-includes its kernel file)."""
    if tilesize not in _BUILT_TILES:
        raise AvsError(-4, "Not supported: tile size %d (built: 32, 64, 128)" % tilesize)
    if tiledim2d and tilesize % tilefactor != 0:
        raise AvsError(-4, "tilefactor must divide tilesize")
    _compiled.update(tilesize=tilesize, tilefactor=tilefactor, tiledim2d=tiledim2d)


def _opts(sched=None, tree=None, domain=0, update_mode=0, workers=0, wait_ms=0, single_stream=False,
          phase_timing=False, reserve_cus=0, graph=False, masked_modes=0):
    o = Opts()
    lib().avs_default_opts(ctypes.byref(o))
    if sched is not None:
        o.sched = sched
    if tree is not None:
        o.tree = tree
    o.domain = domain
    o.update_mode = update_mode
    o.workers = workers
    o.wait_ms = wait_ms
    o.phase_timing = 1 if phase_timing else 0
    o.single_stream = 1 if single_stream else 0
    o.reserve_cus = reserve_cus
    o.graph = 1 if graph else 0
    # update modes on the CU-masked stream (bit 0 GEQRT's, 1 TSQRT's, 2 merges'; 0 = all)
    o.masked_modes = masked_modes
    return o


class TiledMatrix:
    """TiledMatrix(A, blocksize_m, blocksize_n; tiledim2d) (TiledMatrix.jl:41-67).

    Wraps a square column-major host array A (float64 or float32) as an
    nt x nt grid of b x b tiles held tile-major in HBM.  Like the reference it
    aliases A: the tiled algorithms write their result back into A.
    `tiledim2d` is accepted for signature compatibility (the workgroup shape is
    fixed per tile size here).  Extra keyword options select the executor:
    sched (SCHED_STREAMS / SCHED_HOST), tree (TREE_HIER / TREE_FLAT), domain.
    """

    def __init__(self, A, blocksize_m, blocksize_n=None, tiledim2d=0, **opts):
        blocksize_n = blocksize_m if blocksize_n is None else blocksize_n
        if A.ndim != 2 or blocksize_m != blocksize_n or A.shape[0] != A.shape[1] or A.shape[0] % blocksize_n:
            raise AvsError(-3, "Not supported")
        if not A.flags.f_contiguous:
            raise AvsError(-2, "A must be a column-major (Fortran-ordered) array")
        self.A = A
        self.M, self.N = A.shape
        self.m = self.n = blocksize_m
        self.no_tiles = self.M // self.m
        self.dtype = A.dtype
        self._h = ctypes.c_void_p()
        o = _opts(**opts)
        check(lib().avs_create(ctypes.byref(self._h), dtype_code(A.dtype), self.M, self.m, ctypes.byref(o)))
        check(lib().avs_set_matrix_host(self._h, vptr(A), self.M))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().avs_destroy(h)
            self._h = None

    # -- device <-> host
    def reload(self):
        """Re-upload A (after the caller changed it on the host)."""
        check(lib().avs_set_matrix_host(self._h, vptr(self.A), self.M))

    def sync_to_host(self):
        check(lib().avs_get_matrix_host(self._h, vptr(self.A), self.M))

    def band(self):
        """(b+1) x n FP64 LAPACK upper-band array of the reduced matrix."""
        out = np.zeros((self.m + 1, self.M), order="F")
        check(lib().avs_get_band(self._h, dptr(out), self.m + 1))
        return out

    def last_time_ms(self):
        ms = ctypes.c_double()
        check(lib().avs_last_time_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().avs_synchronize(self._h))

    def svdvals(self):
        """Band reduction (if not done) + host bulge chasing + dbdsdc."""
        out = np.empty(self.M)
        check(lib().avs_svdvals(self._h, dptr(out)))
        return out


def apply_q_(A, dC, ldc, ncols, applyt=True):
    """myormqr!(C, A, Tau) (CAQR/src/qr_kernels.jl:256-265): C <- Q^T C (or Q C)
    on the device matrix at address dC (column-major, leading dimension ldc,
    ncols columns, A's dtype), Q from A's last tiled_QR_; returns when done."""
    check(lib().avs_apply_q(A._h, ctypes.c_void_p(dC), int(ldc), int(ncols), 1 if applyt else 0))
    A.synchronize()


def _kend(A, kend):
    return 0 if kend is None else int(kend)


def Blockbidiag_(A, kend=None, sync=True):
    """Blockbidiag!(A::TiledMatrix; kend) (tiledalgos.jl:17-34)."""
    check(lib().avs_blockbidiag(A._h, _kend(A, kend)))
    if sync:
        A.sync_to_host()
    return A


def tiled_QR_(A, kend=None, sync=True):
    """tiled_QR!(A::TiledMatrix; kend) (tiledalgos.jl:6-15)."""
    check(lib().avs_tiled_qr(A._h, _kend(A, kend)))
    if sync:
        A.sync_to_host()
    return A


def _ooc(entry, A, kswitch, mydims2d, tilesize, tilefactor, opts):
    """OOC_alg! (tiledalgos.jl:232-251) on the host matrix A, in place: the
    first n/b - kswitch steps stream A through the GPU (LargeTiledMatrix,
    TiledMatrix.jl:121-186, tiledalgos.jl:75-225), the trailing kswitch tiles
    run in core; n <= kswitch * b runs in core.  mydims2d / tilefactor select
    the reference's launch shape and are accepted for the signature."""
    if A.ndim != 2 or A.shape[0] != A.shape[1] or A.shape[0] % tilesize:
        raise AvsError(-2, "Not supported")
    if not A.flags.f_contiguous:
        raise AvsError(-4, "A must be a column-major (Fortran-ordered) array")
    o = _opts(**opts)
    check(entry(dtype_code(A.dtype), A.shape[0], tilesize, vptr(A), A.shape[0], int(kswitch), ctypes.byref(o)))


def OOC_QR_(A, kswitch=10, mydims2d=False, tilesize=32, tilefactor=1, **opts):
    """OOC_QR! (tiledalgos.jl:253): tiled QR of the host matrix A in place;
    returns nothing, like the reference."""
    _ooc(lib().avs_ooc_tiled_qr, A, kswitch, mydims2d, tilesize, tilefactor, opts)


def OOC_Bidiag_(A, kswitch=10, mydims2d=False, tilesize=32, tilefactor=1, **opts):
    """OOC_Bidiag! (tiledalgos.jl:254): band reduction of the host matrix A in
    place; returns nothing, like the reference."""
    _ooc(lib().avs_ooc_blockbidiag, A, kswitch, mydims2d, tilesize, tilefactor, opts)


def OOC_SVD_(A, kswitch=10, mydims2d=False, tilesize=32, tilefactor=1, **opts):
    """OOC_SVD! (tiledalgos.jl:256-260): band reduction, then bidiagonalize;
    returns the bidiagonal matrix like the reference (diag_lapack gives sigma)."""
    OOC_Bidiag_(A, kswitch, mydims2d, tilesize, tilefactor, **opts)
    return bidiagonalize(A, tilesize)


# --------------------------------------------------------------- host stage
def bidiagonalize_band(band, n, bw):
    """Band (LAPACK upper-band layout, (bw+1) x n FP64) -> (d, e)."""
    band = np.asfortranarray(band, dtype=np.float64)
    d = np.empty(n)
    e = np.empty(max(n - 1, 1))
    check(lib().avs_bidiagonalize_band(n, bw, dptr(band), band.shape[0], dptr(d), dptr(e)))
    return d, e[: n - 1]


def bidiagonalize_band_lapack(band, n, bw):
    """Band -> (d, e) with LAPACK dgbbrd on one thread (the stage-2 path of the
    batched and served forms)."""
    band = np.asfortranarray(band, dtype=np.float64)
    d = np.empty(n)
    e = np.empty(max(n - 1, 1))
    check(lib().avs_bidiagonalize_band_lapack(n, bw, dptr(band), band.shape[0], dptr(d), dptr(e)))
    return d, e[: n - 1]


def bidiagonalize_band_chase(band, n, bw):
    """Band -> (d, e) by Householder bulge chasing, the sweeps pipelined over
    the library's host threads (the default stage 2)."""
    band = np.asfortranarray(band, dtype=np.float64)
    d = np.empty(n)
    e = np.empty(max(n - 1, 1))
    check(lib().avs_bidiagonalize_band_chase(n, bw, dptr(band), band.shape[0], dptr(d), dptr(e)))
    return d, e[: n - 1]


def bidiagonalize_band_device(band, n, bw):
    """Band -> (d, e) by the same bulge chasing on the GPU (bw <= 128): the
    steps of many sweeps at once, one launch per time of the pipelined
    schedule (what svdvals runs from n = 16384 up)."""
    band = np.asfortranarray(band, dtype=np.float64)
    d = np.empty(n)
    e = np.empty(max(n - 1, 1))
    check(lib().avs_bidiagonalize_band_device(n, bw, dptr(band), band.shape[0], dptr(d), dptr(e)))
    return d, e[: n - 1]


def _dense_to_band(A, bw):
    n = A.shape[0]
    band = np.zeros((bw + 1, n), order="F")
    for d in range(bw + 1):
        band[bw - d, d:] = np.diagonal(A, d)
    return band


def bidiagonalize(A, bandwidth):
    """bidiagonalize(A, bandwidth) (bulgechasing.jl:33-42): A (upper band of
    the given bandwidth) is overwritten by its bidiagonal form, in FP64."""
    n = A.shape[0]
    d, e = bidiagonalize_band(_dense_to_band(A, bandwidth), n, bandwidth)
    A[...] = 0
    idx = np.arange(n)
    A[idx, idx] = d
    if n > 1:
        A[idx[:-1], idx[:-1] + 1] = e
    return A


def bdsdc_values(d, e):
    n = len(d)
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e if n > 1 else np.zeros(1), dtype=np.float64)
    out = np.empty(n)
    check(lib().avs_bdsdc_values(n, dptr(d), dptr(e), dptr(out)))
    return out


def diag_lapack(A):
    """diag_lapack(A) (bulgechasing.jl:50-53): dbdsdc('U','N') on diag(A), diag(A,1)."""
    return bdsdc_values(np.diag(A).copy(), np.diag(A, 1).copy())


def my_CU_svdval(A, block_size):
    """my_CU_svdval(A::Matrix, block_size) (bulgechasing.jl:45-48)."""
    return diag_lapack(bidiagonalize(A, block_size))


def svdvals(A, tilesize=None, **opts):
    """Singular values of a square host matrix through the whole path.  Without
    a tile size the autotuner's table picks it (form "svdvals")."""
    A = np.asfortranarray(A)
    if tilesize is None:
        from .autotune import choose_tile
        tilesize, dom = choose_tile(A.shape[0], A.dtype, "svdvals")
        if dom and "domain" not in opts:
            opts["domain"] = dom
    return TiledMatrix(A, tilesize, tilesize, **opts).svdvals()


def gaussian_matrix(n, seed, scale=1.0, dtype=np.float64):
    """Seeded N(0, scale^2) matrix (Philox4x32-10 + Box-Muller), column-major."""
    A = np.empty((n, n), dtype=dtype, order="F")
    check(lib().avs_gen_gaussian_host(seed, n, scale, dtype_code(dtype), vptr(A), n))
    return A


# --------------------------------------------------------- tile operators
# Device tensors are passed as (pointer, leading dimension); torch CUDA tensors
# (column-major views, i.e. t.T of a row-major tensor) are the usual carrier.
def _dev(t):
    return ctypes.c_void_p(t.data_ptr())


def geqrt_(dtype, b, dA, lda, dtau, trans=False, stream=None):
    """QR_unsafe_kernel! (QRkernels.jl:3-66) on a device tile."""
    check(lib().avs_geqrt(dtype, b, dA, lda, dtau, int(trans), stream))


def tsqrt_(dtype, b, dA1, lda1, dA2, lda2, dtau, trans=False, stream=None):
    """QR_unsafe_kernel2! (QRkernels.jl:71-129)."""
    check(lib().avs_tsqrt(dtype, b, dA1, lda1, dA2, lda2, dtau, int(trans), stream))


def unmqr_(dtype, b, dV, ldv, dtau, dC, ldc, ncols, applyt=True, vtrans=False, ctrans=False, stream=None):
    """applyQorQt_unsafe_kernel! (QRkernels.jl:131-161)."""
    check(lib().avs_unmqr(dtype, b, dV, ldv, dtau, dC, ldc, ncols, int(applyt), int(vtrans), int(ctrans), stream))


def tsmqr_(dtype, b, dC1, ldc1, dC2, ldc2, dV, ldv, dtau, ncols, applyt=True, vtrans=False, ctrans=False,
           stream=None):
    """applyQorQt_unsafe_kernel2! (QRkernels.jl:165-197)."""
    check(lib().avs_tsmqr(dtype, b, dC1, ldc1, dC2, ldc2, dV, ldv, dtau, ncols, int(applyt), int(vtrans),
                          int(ctrans), stream))


def schedule_stats(nt, b, alg=0, kend=0, tree=1, domain=16, workers=256):
    n = ctypes.c_int64()
    est = ctypes.c_double()
    check(lib().avs_schedule_stats(nt, b, alg, kend, tree, domain, workers, ctypes.byref(n), ctypes.byref(est)))
    return n.value, est.value


def schedule_dump(nt, b, alg=0, kend=0, tree=1, domain=16, workers=256):
    """Task list of the persistent executor in dispatch order (list of dicts)."""
    need = ctypes.c_int64()
    check(lib().avs_schedule_dump(nt, b, alg, kend, tree, domain, workers, None, 0, ctypes.byref(need)))
    buf = np.zeros(max(need.value, 1), dtype=np.int32)
    check(lib().avs_schedule_dump(nt, b, alg, kend, tree, domain, workers,
                                  buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), need.value,
                                  ctypes.byref(need)))
    out, i = [], 0
    v = buf.tolist()
    keys = ("kind", "mode", "trans", "ntiles", "vtile", "btile", "c1tile", "cstep", "fac")
    while i < need.value:
        t = dict(zip(keys, v[i:i + 9]))
        nd = v[i + 9]
        i += 10
        t["deps"] = [(v[i + 2 * d], v[i + 2 * d + 1]) for d in range(nd)]
        i += 2 * nd
        ns = v[i]
        t["sigs"] = v[i + 1:i + 1 + ns]
        i += 1 + ns
        out.append(t)
    return out


def phase_dump(nt, b, alg=0, kend=0, tree=1, domain=16):
    """Phases of the stream-ordered executor: list of (kind, ncols, [chain, ...]),
    each chain a list of op dicts."""
    need = ctypes.c_int64()
    check(lib().avs_phase_dump(nt, b, alg, kend, tree, domain, None, 0, ctypes.byref(need)))
    buf = np.zeros(max(need.value, 1), dtype=np.int32)
    check(lib().avs_phase_dump(nt, b, alg, kend, tree, domain, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               need.value, ctypes.byref(need)))
    v = buf.tolist()
    keys = ("kind", "mode", "trans", "vtile", "btile", "c1tile", "cstep", "fac", "ncols")
    i, phases = 1, []
    for _ in range(v[0]):
        kind, ncols, nch = v[i], v[i + 1], v[i + 2]
        i += 3
        chains = []
        for _ in range(nch):
            nops = v[i]
            i += 1
            chains.append([dict(zip(keys, v[i + 9 * q:i + 9 * q + 9])) for q in range(nops)])
            i += 9 * nops
        phases.append((kind, ncols, chains))
    return phases


def mfma_layout(variant=0):
    o64 = np.zeros(256)
    o32 = np.zeros(256, dtype=np.float32)
    check(lib().avs_debug_mfma_layout(variant, dptr(o64), o32.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return o64.reshape(64, 4), o32.reshape(64, 4)
