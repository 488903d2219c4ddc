"""avs_tsmqr_noalloc, the TSMQR that allocates nothing, against avs_tsmqr.

Both run the same kernels (T_JJ build, then the compact-WY update of
applyQorQt_unsafe_kernel2!, KAbasedSVD/src/QRkernels.jl:165-197); the only
difference is where T_JJ lives (a fixed per-GPU array instead of a hipMalloc'd
buffer), so the results must agree bit for bit.  Two calls in a row share the
fixed array; the second has the smaller input, so a stale T_JJ from the first
call would show.
"""
import ctypes

import numpy as np
import pytest

import avsvd
from avsvd._lib import check, lib

pytestmark = pytest.mark.gpu

DT = {np.float64: avsvd.AVS_F64, np.float32: avsvd.AVS_F32}


def _dev(a, dt, gpu):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dt).T)).to(gpu)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _factor(gpu, dt, b, rng):
    """TSQRT reflectors (V in the bottom tile) and tau of a random tile pair."""
    import torch
    A1 = _dev(rng.standard_normal((b, b)), dt, gpu)
    A2 = _dev(rng.standard_normal((b, b)), dt, gpu)
    tau = torch.zeros(b, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    avsvd.tsqrt_(DT[dt], b, _ptr(A1), b, _ptr(A2), b, _ptr(tau))
    torch.cuda.synchronize()
    return A2, tau


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_tsmqr_noalloc_matches_tsmqr_twice(gpu, dt):
    import torch
    rng = np.random.default_rng(20261204)
    side = torch.cuda.Stream(device=gpu)
    # (b, column tiles, applyt, ctrans): the second call is the smaller one
    for call, (b, ntl, applyt, ctrans) in enumerate(((128, 4, True, False), (64, 2, False, True))):
        print("no-alloc TSMQR call %d: dtype %s b=%d ncols=%d applyt=%d ctrans=%d"
              % (call, np.dtype(dt).name, b, ntl * b, applyt, ctrans), flush=True)
        V, tau = _factor(gpu, dt, b, rng)
        ncols = ntl * b
        shape = (ncols, b) if ctrans else (b, ncols)
        ldc = shape[0]
        c1 = rng.standard_normal(shape)
        c2 = rng.standard_normal(shape)
        R1, R2 = _dev(c1, dt, gpu), _dev(c2, dt, gpu)
        N1, N2 = _dev(c1, dt, gpu), _dev(c2, dt, gpu)
        torch.cuda.synchronize()
        avsvd.tsmqr_(DT[dt], b, _ptr(R1), ldc, _ptr(R2), ldc, _ptr(V), b, _ptr(tau), ncols, applyt=applyt,
                     ctrans=ctrans)
        check(lib().avs_tsmqr_noalloc(DT[dt], b, _ptr(N1), ldc, _ptr(N2), ldc, _ptr(V), b, _ptr(tau), ncols,
                                      int(applyt), 0, int(ctrans), ctypes.c_void_p(side.cuda_stream)))
        side.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(N1, R1), "C1 differs from avs_tsmqr (call %d)" % call
        assert torch.equal(N2, R2), "C2 differs from avs_tsmqr (call %d)" % call
        # the update did something: C changed from its input
        assert not torch.equal(N2, _dev(c2, dt, gpu))
