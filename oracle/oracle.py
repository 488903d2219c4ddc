"""Python handle on the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker or the timed CPU baseline, never by the shipped package.

Every function mirrors the reference function named in its docstring
(paths relative to the reference repository root).  LAPACK routines used as
independent pins come from the OpenBLAS that scipy bundles (scipy_* symbols).
"""
import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LAPACK = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i, l, d = ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.or_geqrt.argtypes = [i, _dp, l, i, _dp]
        L.or_tsqrt.argtypes = [i, _dp, l, _dp, l, i, _dp]
        L.or_unmqr.argtypes = [i, _dp, l, i, _dp, _dp, l, i, l, i]
        L.or_tsmqr.argtypes = [i, _dp, l, _dp, l, i, _dp, l, i, _dp, l, i]
        L.or_blockbidiag.argtypes = [l, i, _dp, i]
        L.or_tiled_qr.argtypes = [l, i, _dp, i]
        L.or_bidiagonalize.argtypes = [l, l, _dp]
        L.or_block_bidiagonalize.argtypes = [_dp, l, l, l]
        L.or_philox_normal.argtypes = [ctypes.c_uint64, l, l, d, _dp]
        L.cp_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.cp_init.restype = ctypes.c_int
        L.cp_blas_threads.argtypes = [i]
        L.cp_blas_threads.restype = ctypes.c_int
        L.cp_band_bidiagonal.argtypes = [l, i, _dp, i]
        L.cp_band_bidiagonal.restype = ctypes.c_int
        for f in ("or_geqrt", "or_tsqrt", "or_unmqr", "or_tsmqr", "or_blockbidiag",
                  "or_tiled_qr", "or_bidiagonalize", "or_block_bidiagonalize",
                  "or_philox_normal"):
            getattr(L, f).restype = None
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def _fortran(a):
    return np.asfortranarray(np.array(a, dtype=np.float64, copy=True))


# --------------------------------------------------------------------------
# tile kernels (KAbasedSVD/src/QRkernels.jl)
# --------------------------------------------------------------------------
def geqrt(a, trans=False):
    """QR_unsafe_kernel!, QRkernels.jl:3-66. Returns (factors, tau[b])."""
    a = _fortran(a)
    b = a.shape[0]
    tau = np.zeros(b)
    lib().or_geqrt(b, _p(a), b, int(trans), _p(tau))
    return a, tau


def tsqrt(a1, a2, trans=False):
    """QR_unsafe_kernel2!, QRkernels.jl:71-129. Returns (a1', a2' = V2, tau)."""
    a1, a2 = _fortran(a1), _fortran(a2)
    b = a1.shape[0]
    tau = np.zeros(b)
    lib().or_tsqrt(b, _p(a1), b, _p(a2), b, int(trans), _p(tau))
    return a1, a2, tau


def unmqr(v, tau, c, applyt=True, vtrans=False, ctrans=False):
    """applyQorQt_unsafe_kernel!, QRkernels.jl:131-161 on a b x ncols block
    (or its adjoint when ctrans, as the reference's `bcopy'` calls)."""
    v, c = _fortran(v), _fortran(c)
    b = v.shape[0]
    ncols = c.shape[0] if ctrans else c.shape[1]
    lib().or_unmqr(b, _p(v), b, int(vtrans), _p(np.ascontiguousarray(tau, dtype=np.float64)),
                   _p(c), c.shape[0], int(ctrans), ncols, int(applyt))
    return c


def tsmqr(c1, c2, v, tau, applyt=True, ctrans=False, vtrans=False):
    """applyQorQt_unsafe_kernel2!, QRkernels.jl:165-197."""
    c1, c2, v = _fortran(c1), _fortran(c2), _fortran(v)
    b = v.shape[0]
    ncols = c1.shape[0] if ctrans else c1.shape[1]
    lib().or_tsmqr(b, _p(c1), c1.shape[0], _p(c2), c2.shape[0], int(ctrans), _p(v), b,
                   int(vtrans), _p(np.ascontiguousarray(tau, dtype=np.float64)), ncols,
                   int(applyt))
    return c1, c2


# --------------------------------------------------------------------------
# tiled algorithms (KAbasedSVD/src/tiledalgos.jl)
# --------------------------------------------------------------------------
def blockbidiag(a, b, kend=None):
    """Blockbidiag!, tiledalgos.jl:17-34; returns the reduced n x n matrix."""
    a = _fortran(a)
    n = a.shape[0]
    assert a.shape == (n, n) and n % b == 0
    kend = n // b if kend is None else kend
    lib().or_blockbidiag(n, b, _p(a), kend)
    return a


def tiled_qr(a, b, kend=None):
    """tiled_QR!, tiledalgos.jl:6-15."""
    a = _fortran(a)
    n = a.shape[0]
    kend = n // b if kend is None else kend
    lib().or_tiled_qr(n, b, _p(a), kend)
    return a


def bidiagonalize(a, bw):
    """bidiagonalize(A, bandwidth), bulgechasing.jl:33-42 (dense, in place)."""
    a = _fortran(a)
    lib().or_bidiagonalize(a.shape[0], bw, _p(a))
    return a


def philox_normal(seed, count, scale=1.0, first=0):
    out = np.empty(count)
    lib().or_philox_normal(seed, first, count, scale, _p(out))
    return out


def gaussian_matrix(n, seed, scale=1.0):
    """n x n column-major matrix with iid N(0, scale^2) entries from (seed)."""
    return philox_normal(seed, n * n, scale).reshape((n, n), order="F")


# --------------------------------------------------------------------------
# LAPACK from the OpenBLAS bundled with scipy (independent pins)
# --------------------------------------------------------------------------
def lapack():
    global _LAPACK
    if _LAPACK is None:
        import scipy
        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
        if not cands:
            raise RuntimeError("no OpenBLAS/LAPACK found next to scipy")
        _LAPACK = ctypes.CDLL(cands[0])
    return _LAPACK


def bdsdc(d, e):
    """LAPACK.bdsdc!('U','N', d, e) as diag_lapack does (bulgechasing.jl:50-53);
    descending singular values of the upper bidiagonal (d, e)."""
    L = lapack()
    n = len(d)
    dd = np.array(d, dtype=np.float64, copy=True)
    ee = np.zeros(max(n, 1), dtype=np.float64)
    ee[: n - 1] = np.asarray(e, dtype=np.float64)[: n - 1]
    work = np.zeros(max(1, 4 * n))
    iwork = np.zeros(8 * max(n, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    ci = ctypes.c_int
    one = ci(1)
    dummy = np.zeros(1)
    idummy = np.zeros(1, dtype=np.int32)
    fn = L.scipy_dbdsdc_
    fn.restype = None
    fn(ctypes.c_char_p(b"U"), ctypes.c_char_p(b"N"), ctypes.byref(ci(n)), _p(dd), _p(ee),
       _p(dummy), ctypes.byref(one), _p(dummy), ctypes.byref(one), _p(dummy),
       idummy.ctypes.data_as(ctypes.POINTER(ci)), _p(work),
       iwork.ctypes.data_as(ctypes.POINTER(ci)), ctypes.byref(info),
       ctypes.c_size_t(1), ctypes.c_size_t(1))
    if info.value != 0:
        raise RuntimeError("dbdsdc info=%d" % info.value)
    return dd


_CP_READY = False


def _cp_ready():
    global _CP_READY
    if not _CP_READY:
        lapack()
        st = lib().cp_init(_LAPACK._name.encode(), b"scipy_")
        if st:
            raise RuntimeError("cpu_plain: LAPACK dgeqrt/dgemqrt not found (%d)" % st)
        _CP_READY = True


def band_bidiagonal_cpu(a, b, nthreads=1):
    """BandBidiagonal!(A, b, b, nt, nt, false), CUDAbasedSVD/SVD_GPU.jl:67-214:
    the reference's CPU plain path (LAPACK QR per cached 2b x b block, Q
    applied to the other blocks; the block loop over `nthreads` threads with a
    single-threaded BLAS).  Returns the reduced n x n matrix."""
    _cp_ready()
    a = _fortran(a)
    n = a.shape[0]
    assert a.shape == (n, n) and n % b == 0
    prev = lib().cp_blas_threads(1)
    try:
        st = lib().cp_band_bidiagonal(n, b, _p(a), nthreads)
    finally:
        if prev > 0:
            lib().cp_blas_threads(prev)
    if st:
        raise RuntimeError("cp_band_bidiagonal status %d" % st)
    return a


def svdvals_host_stage(band_matrix, bw):
    """bidiagonalize + diag_lapack (bulgechasing.jl:33-53): sigma of a band."""
    a = bidiagonalize(band_matrix, bw)
    return bdsdc(np.diag(a).copy(), np.diag(a, 1).copy())


def svdvals_ref(a):
    """FP64 LAPACK gesdd singular values (the end-to-end oracle)."""
    import scipy.linalg
    return scipy.linalg.svd(np.asarray(a, dtype=np.float64), compute_uv=False, lapack_driver="gesdd")


def sigma_err(sig, ref):
    """max_i |sigma_i - ref_i| / ref_max, both sorted descending."""
    s = np.sort(np.asarray(sig, dtype=np.float64))[::-1]
    r = np.sort(np.asarray(ref, dtype=np.float64))[::-1]
    return float(np.max(np.abs(s - r)) / r[0])
