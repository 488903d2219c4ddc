"""Out-of-core runs (OOC_Bidiag! / OOC_QR! / OOC_SVD!, KAbasedSVD/src/
tiledalgos.jl:232-260) against the in-core run of the same algorithm.

The streamed steps apply the in-core flat-tree executors' kernels in the same
per-tile order, so with the flat tree everywhere the host result must equal
the in-core result bit for bit: every tile row and column that was streamed
(band with exact zeros outside it, or R and reflectors), and the trailing
kswitch block that ran in core.  Sizes follow the reference's OOC tests
(KAbasedSVD/test/test_tiledalgo.jl runs OOC_* at kswitch < n/b).
"""
import ctypes

import numpy as np
import pytest

import avsvd
from avsvd._lib import lib

pytestmark = pytest.mark.gpu


def _matrix(n, dt, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, n)).astype(dt))


def _in_core(A, b, alg):
    T = avsvd.TiledMatrix(A, b, tree=avsvd.TREE_FLAT)
    (avsvd.Blockbidiag_ if alg == "bidiag" else avsvd.tiled_QR_)(T)
    del T
    return A


@pytest.mark.parametrize("dt,n,b,kswitch", [
    (np.float64, 1024, 32, 8),
    (np.float64, 1024, 64, 3),
    (np.float64, 1024, 128, 2),
    (np.float32, 1024, 32, 8),
    (np.float64, 512, 32, 0),    # every step streamed
    (np.float64, 512, 32, 16),   # n == kswitch * b: all in core
])
def test_ooc_bidiag_equals_in_core_flat_bitwise(gpu, dt, n, b, kswitch):
    A0 = _matrix(n, dt, 1000 + n + b + kswitch)
    ref = _in_core(A0.copy(order="F"), b, "bidiag")
    A = A0.copy(order="F")
    assert avsvd.OOC_Bidiag_(A, kswitch=kswitch, tilesize=b, tree=avsvd.TREE_FLAT) is None
    np.testing.assert_array_equal(A, ref)
    # the result is a band of bandwidth b
    np.testing.assert_array_equal(np.tril(A, -1), 0)
    np.testing.assert_array_equal(np.triu(A, b + 1), 0)


@pytest.mark.parametrize("dt,n,b,kswitch", [(np.float64, 1024, 32, 8), (np.float32, 512, 64, 2)])
def test_ooc_tiled_qr_equals_in_core_flat_bitwise(gpu, dt, n, b, kswitch):
    A0 = _matrix(n, dt, 2000 + n + b)
    ref = _in_core(A0.copy(order="F"), b, "qr")
    A = A0.copy(order="F")
    assert avsvd.OOC_QR_(A, kswitch=kswitch, tilesize=b, tree=avsvd.TREE_FLAT) is None
    np.testing.assert_array_equal(A, ref)
    # |R| matches LAPACK's up to the signs of its rows
    R = np.triu(A).astype(np.float64)
    Rl = np.linalg.qr(A0.astype(np.float64), mode="r")
    tol = 1e-10 if dt == np.float64 else 2e-3
    np.testing.assert_allclose(np.abs(R), np.abs(Rl), atol=tol * np.abs(Rl).max())


def test_ooc_bidiag_with_leading_dimension(gpu):
    n, b, kswitch, ld = 512, 32, 4, 544
    A0 = _matrix(n, np.float64, 3001)
    ref = _in_core(A0.copy(order="F"), b, "bidiag")
    big = np.asfortranarray(np.full((ld, n), 7.0))
    big[:n, :] = A0
    o = avsvd.tiled._opts(tree=avsvd.TREE_FLAT)
    st = lib().avs_ooc_blockbidiag(avsvd.AVS_F64, n, b, big.ctypes.data_as(ctypes.c_void_p), ld, kswitch,
                                   ctypes.byref(o))
    assert st == 0, lib().avs_last_error()
    np.testing.assert_array_equal(big[:n, :], ref)
    np.testing.assert_array_equal(big[n:, :], 7.0)   # rows past n untouched


def test_ooc_svd_sigma_matches_lapack(gpu):
    n, b = 512, 32
    A0 = _matrix(n, np.float64, 4001)
    A = A0.copy(order="F")
    bid = avsvd.OOC_SVD_(A, kswitch=4, tilesize=b)   # trailing block on the default (hierarchical) tree
    s = avsvd.diag_lapack(bid)
    ref = np.linalg.svd(A0, compute_uv=False)
    assert np.max(np.abs(np.sort(s)[::-1] - ref)) <= 1e-12 * ref[0]
