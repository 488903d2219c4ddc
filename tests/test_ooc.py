"""Out-of-core entry points on the CPU: argument errors (raised before any
device call) and the return conventions of the OOC_* wrappers.

OOC_QR! / OOC_Bidiag! (KAbasedSVD/src/tiledalgos.jl:253-254) reduce their
host matrix in place and return nothing; OOC_SVD! (:256-260) returns
bidiagonalize(A, tilesize), the matrix overwritten by its bidiagonal form.
"""
import ctypes

import numpy as np
import pytest

import avsvd
from avsvd import tiled
from avsvd._lib import lib


def test_ooc_argument_errors_before_any_device_call():
    L = lib()
    A = np.zeros((64, 64), order="F")
    p = A.ctypes.data_as(ctypes.c_void_p)
    for entry in (L.avs_ooc_blockbidiag, L.avs_ooc_tiled_qr):
        assert entry(7, 64, 32, p, 64, 1, None) == -1
        assert entry(0, 100, 32, p, 100, 1, None) == -2
        assert b"Not supported" in L.avs_last_error()
        assert entry(0, 96, 48, p, 96, 1, None) == -3
        assert entry(0, 64, 32, None, 64, 1, None) == -4
        assert entry(0, 64, 32, p, 32, 1, None) == -5
        assert entry(0, 64, 32, p, 64, -1, None) == -6
    with pytest.raises(avsvd.AvsError):
        avsvd.OOC_Bidiag_(np.zeros((100, 100), order="F"), kswitch=2, tilesize=32)
    with pytest.raises(avsvd.AvsError):
        avsvd.OOC_QR_(np.zeros((64, 64)), kswitch=2, tilesize=32)  # C order


class _FakeLib:
    """Stands in for the device entry points: records the call and leaves an
    upper band of bandwidth b in A, as the band reduction would."""

    def __init__(self, real, band):
        self.real, self.band, self.calls = real, band, []

    def _run(self, name, dtype, n, b, pA, lda, kswitch, opts):
        self.calls.append((name, dtype, n, b, lda, kswitch))
        A = np.ctypeslib.as_array(ctypes.cast(pA, ctypes.POINTER(ctypes.c_double)), shape=(n * lda,))
        A[:] = self.band.reshape(-1, order="F")
        return 0

    def avs_ooc_blockbidiag(self, *a):
        return self._run("bidiag", *a)

    def avs_ooc_tiled_qr(self, *a):
        return self._run("qr", *a)

    def __getattr__(self, name):
        return getattr(self.real, name)


def test_ooc_wrappers_keep_the_reference_return_conventions(monkeypatch):
    n, b = 96, 32
    rng = np.random.default_rng(7)
    band = np.triu(np.tril(rng.standard_normal((n, n)), b))
    fake = _FakeLib(lib(), band)
    monkeypatch.setattr(tiled, "lib", lambda: fake)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    assert avsvd.OOC_Bidiag_(A, kswitch=1, tilesize=b) is None
    np.testing.assert_array_equal(A, band)           # in place
    assert avsvd.OOC_QR_(A, kswitch=2, tilesize=b) is None
    assert [c[0] for c in fake.calls] == ["bidiag", "qr"]
    assert fake.calls[0][1:] == (0, n, b, n, 1)      # dtype, n, b, lda, kswitch reach the C ABI
    B = np.asfortranarray(rng.standard_normal((n, n)))
    out = avsvd.OOC_SVD_(B, kswitch=3, tilesize=b)
    assert out is B                                  # bidiagonalize returns the matrix it overwrote
    assert fake.calls[-1][0] == "bidiag" and fake.calls[-1][5] == 3
    np.testing.assert_array_equal(np.tril(out, -1), 0)
    np.testing.assert_array_equal(np.triu(out, 2), 0)
    # the bidiagonal keeps the band's singular values
    np.testing.assert_allclose(np.linalg.svd(out, compute_uv=False), np.linalg.svd(band, compute_uv=False),
                               rtol=1e-12, atol=1e-12 * np.abs(band).max())
