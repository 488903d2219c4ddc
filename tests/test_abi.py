"""The C ABI library loads, exports every symbol include/avsvd.h declares, and
turns bad arguments into the documented statuses (CPU: no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import avsvd
import oracle as O
from avsvd._lib import Opts, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "avsvd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(avs_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(avsvd.EXPORTS)


def test_create_rejects_unsupported_shapes_like_the_reference():
    # TiledMatrix.jl:42-44 raises "Not supported" when n % b != 0
    L = lib()
    h = ctypes.c_void_p()
    o = Opts()
    L.avs_default_opts(ctypes.byref(o))
    assert L.avs_create(ctypes.byref(h), 0, 100, 32, ctypes.byref(o)) == -3
    assert b"Not supported" in L.avs_last_error()
    assert L.avs_create(ctypes.byref(h), 0, 96, 48, ctypes.byref(o)) == -4
    assert L.avs_create(ctypes.byref(h), 7, 64, 32, ctypes.byref(o)) == -2
    with pytest.raises(avsvd.AvsError):
        avsvd.TiledMatrix(np.zeros((100, 100), order="F"), 32, 32)
    with pytest.raises(avsvd.AvsError):
        avsvd.compile_kernels(tilesize=48)


def test_tsmqr_noalloc_refuses_b_beyond_its_workspace():
    # the fixed T_JJ workspace holds 16 * 128 doubles; the check runs before any HIP call
    L = lib()
    nul = ctypes.c_void_p()
    assert L.avs_tsmqr_noalloc(0, 256, nul, 256, nul, 256, nul, 256, nul, 256, 1, 0, 0, nul) == -2
    assert b"workspace" in L.avs_last_error()


def test_generator_matches_oracle_bitwise():
    for n, seed in ((5, 1), (64, 20261203), (129, 77)):
        np.testing.assert_array_equal(avsvd.gaussian_matrix(n, seed), O.gaussian_matrix(n, seed))
    a32 = avsvd.gaussian_matrix(33, 9, dtype=np.float32)
    np.testing.assert_array_equal(a32, O.gaussian_matrix(33, 9).astype(np.float32))


@pytest.mark.parametrize("n,bw", [(64, 8), (128, 32), (192, 64), (512, 128)])
def test_host_stage_matches_oracle_bitwise(n, bw):
    """bidiagonalize (bulgechasing.jl:1-42) in the library's band store equals
    the oracle's dense restatement bit for bit; bdsdc gives sigma(band)."""
    import scipy.linalg as sl
    a = O.gaussian_matrix(n, 20261203 + 13 * n + bw)
    band = np.triu(np.tril(a, bw))
    d, e = avsvd.bidiagonalize_band(avsvd.tiled._dense_to_band(band, bw), n, bw)
    od = O.bidiagonalize(band, bw)
    np.testing.assert_array_equal(d, np.diag(od))
    np.testing.assert_array_equal(e, np.diag(od, 1))
    sig = avsvd.bdsdc_values(d, e)
    assert O.sigma_err(sig, sl.svdvals(band)) < 10 * n * np.finfo(float).eps
    np.testing.assert_array_equal(sig, O.bdsdc(d, e))


@pytest.mark.parametrize("n,bw", [(1, 4), (5, 4), (64, 8), (300, 32), (512, 64)])
def test_lapack_stage2_preserves_sigma(n, bw):
    """The one-thread stage 2 of the batched and served forms (LAPACK dgbbrd):
    sigma(d, e) = sigma(band) to FP64 rounding, and = sigma of the reference's
    block sweeps on the same band."""
    import scipy.linalg as sl
    a = O.gaussian_matrix(max(n, 2), 20261211 + n + bw)[:n, :n]
    band = np.triu(np.tril(a, bw))
    st = avsvd.tiled._dense_to_band(band, bw)
    d, e = avsvd.tiled.bidiagonalize_band_lapack(st, n, bw)
    sig = avsvd.bdsdc_values(d, e)
    ref = sl.svdvals(band) if n > 1 else np.abs(np.diag(band))
    assert O.sigma_err(sig, ref) < 10 * max(n, 2) * np.finfo(float).eps
    dh, eh = avsvd.bidiagonalize_band(st, n, bw)
    assert O.sigma_err(sig, avsvd.bdsdc_values(dh, eh)) < 10 * max(n, 2) * np.finfo(float).eps


@pytest.mark.parametrize("n,bw", [(1, 4), (2, 4), (5, 4), (7, 3), (64, 8), (100, 1), (257, 31), (300, 32),
                                  (512, 64), (700, 16)])
def test_chase_stage2_preserves_sigma(n, bw):
    """The Householder bulge chasing (default stage 2 of the batched, served and
    plain forms): sigma(d, e) = sigma(band) to FP64 rounding, the same as
    LAPACK dgbbrd's, and bit-identical for 1 and 4 pipelined host threads."""
    import scipy.linalg as sl
    a = O.gaussian_matrix(max(n, 2), 20261231 + n + bw)[:n, :n]
    band = np.triu(np.tril(a, bw))
    st = avsvd.tiled._dense_to_band(band, bw)
    outs = []
    old = os.environ.get("AVSVD_HOST_THREADS")
    try:
        for th in ("1", "4"):
            os.environ["AVSVD_HOST_THREADS"] = th
            outs.append(avsvd.tiled.bidiagonalize_band_chase(st, n, bw))
    finally:
        if old is None:
            os.environ.pop("AVSVD_HOST_THREADS", None)
        else:
            os.environ["AVSVD_HOST_THREADS"] = old
    (d, e), (d4, e4) = outs
    np.testing.assert_array_equal(d, d4)
    np.testing.assert_array_equal(e, e4)
    sig = avsvd.bdsdc_values(d, e)
    ref = sl.svdvals(band) if n > 1 else np.abs(np.diag(band))
    assert O.sigma_err(sig, ref) < 10 * max(n, 2) * np.finfo(float).eps
    dl, el = avsvd.tiled.bidiagonalize_band_lapack(st, n, bw)
    assert O.sigma_err(sig, avsvd.bdsdc_values(dl, el)) < 10 * max(n, 2) * np.finfo(float).eps


def test_chase_stage2_graded_band():
    """A band whose entries span 1e-300 .. 1: the chased bulges decay to
    subnormal sizes without producing Inf or NaN (divisions, not reciprocals)."""
    import scipy.linalg as sl
    n, bw = 600, 16
    a = O.gaussian_matrix(n, 20261237)
    g = np.logspace(0, -300, n)
    band = np.triu(np.tril(a * g[:, None], bw))
    st = avsvd.tiled._dense_to_band(band, bw)
    d, e = avsvd.tiled.bidiagonalize_band_chase(st, n, bw)
    assert np.isfinite(d).all() and np.isfinite(e).all()
    sig = avsvd.bdsdc_values(d, e)
    assert O.sigma_err(sig, sl.svdvals(band)) < 10 * n * np.finfo(float).eps


def test_host_stage_argument_errors():
    L = lib()
    d = np.zeros(4)
    assert L.avs_bidiagonalize_band(0, 4, avsvd._lib.dptr(d), 5, avsvd._lib.dptr(d), avsvd._lib.dptr(d)) == -1
    assert L.avs_bidiagonalize_band(4, 4, avsvd._lib.dptr(d), 3, avsvd._lib.dptr(d), avsvd._lib.dptr(d)) == -4
    assert L.avs_bdsdc_values(4, avsvd._lib.dptr(d), avsvd._lib.dptr(d), None) == -4
    assert L.avs_bidiagonalize_band_lapack(4, 4, avsvd._lib.dptr(d), 3, avsvd._lib.dptr(d),
                                           avsvd._lib.dptr(d)) == -4
    assert L.avs_bidiagonalize_band_chase(4, 4, avsvd._lib.dptr(d), 3, avsvd._lib.dptr(d),
                                          avsvd._lib.dptr(d)) == -4
    assert L.avs_bidiagonalize_band_chase(4, 0, avsvd._lib.dptr(d), 5, avsvd._lib.dptr(d),
                                          avsvd._lib.dptr(d)) == -2
    assert L.avs_bidiagonalize_band_chase(4, 4, avsvd._lib.dptr(d), 5, avsvd._lib.dptr(d), None) == -6
    # the device chase checks its arguments before any device call
    assert L.avs_bidiagonalize_band_device(0, 4, avsvd._lib.dptr(d), 5, avsvd._lib.dptr(d),
                                           avsvd._lib.dptr(d)) == -1
    assert L.avs_bidiagonalize_band_device(4, 129, avsvd._lib.dptr(d), 130, avsvd._lib.dptr(d),
                                           avsvd._lib.dptr(d)) == -2
    assert L.avs_bidiagonalize_band_device(4, 4, avsvd._lib.dptr(d), 3, avsvd._lib.dptr(d),
                                           avsvd._lib.dptr(d)) == -4
    assert L.avs_bidiagonalize_band_device(4, 4, avsvd._lib.dptr(d), 5, avsvd._lib.dptr(d), None) == -6


def test_reference_api_mirror_on_host():
    # my_CU_svdval / diag_lapack / bidiagonalize on a dense band (bulgechasing.jl:33-53)
    import scipy.linalg as sl
    n, bw = 96, 16
    a = O.gaussian_matrix(n, 5)
    band = np.asfortranarray(np.triu(np.tril(a, bw)))
    sig = avsvd.my_CU_svdval(band.copy(order="F"), bw)
    assert O.sigma_err(sig, sl.svdvals(band)) < 10 * n * np.finfo(float).eps
