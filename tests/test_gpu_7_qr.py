"""tiled_QR! and its Q applied to other matrices (myormqr!, CAQR/src/
qr_kernels.jl:256-265) against FP64 LAPACK.

A = Q R with R unique up to the signs of its rows, so Q_ours = Q_lapack D with
D diagonal +-1 and Q_ours^T C = D Q_lapack^T C: the comparison with LAPACK is
on |Q^T C| elementwise (as CAQR/src/benchmark.jl compares with the vendor
ormqr).  Two identities need no reference: Q^T A = R (the factorization's own
upper triangle) and Q (Q^T C) = C.
"""
import ctypes

import numpy as np
import pytest

import avsvd
from avsvd._lib import lib

pytestmark = pytest.mark.gpu


def _dev(a, dt, gpu):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dt).T)).to(gpu)


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().T.astype(np.float64)


@pytest.mark.parametrize("dt,n,b,tree", [
    (np.float64, 1024, 32, avsvd.TREE_HIER),
    (np.float64, 1024, 128, avsvd.TREE_FLAT),
    (np.float64, 512, 64, avsvd.TREE_HIER),
    (np.float32, 1024, 64, avsvd.TREE_HIER),
])
def test_apply_q_vs_lapack(gpu, dt, n, b, tree):
    rng = np.random.default_rng(n + b + tree)
    A0 = np.asfortranarray(rng.standard_normal((n, n)).astype(dt))
    m = 3 * b
    C0 = rng.standard_normal((n, m)).astype(dt)
    A = A0.copy(order="F")
    T = avsvd.TiledMatrix(A, b, tree=tree)
    avsvd.tiled_QR_(T)
    R = np.triu(A.astype(np.float64))
    eps = np.finfo(dt).eps
    # Q^T A0 = R
    dA = _dev(A0, dt, gpu)
    avsvd.apply_q_(T, dA.data_ptr(), n, n, applyt=True)
    QtA = _host(dA)
    assert np.linalg.norm(QtA - R) <= 50 * n * eps * np.linalg.norm(A0)
    # |Q^T C| against LAPACK
    dC = _dev(C0, dt, gpu)
    avsvd.apply_q_(T, dC.data_ptr(), n, m, applyt=True)
    QtC = _host(dC)
    Ql, _ = np.linalg.qr(A0.astype(np.float64), mode="complete")
    ref = Ql.T @ C0.astype(np.float64)
    assert np.linalg.norm(np.abs(QtC) - np.abs(ref)) <= 50 * n * eps * np.linalg.norm(ref)
    # Q (Q^T C) = C
    avsvd.apply_q_(T, dC.data_ptr(), n, m, applyt=False)
    np.testing.assert_allclose(_host(dC), C0.astype(np.float64), atol=50 * n * eps * np.abs(C0).max())


def test_apply_q_needs_tiled_qr_factors(gpu):
    n, b = 256, 32
    A = np.asfortranarray(np.random.default_rng(1).standard_normal((n, n)))
    T = avsvd.TiledMatrix(A, b)
    dC = _dev(np.zeros((n, b)), np.float64, gpu)
    st = lib().avs_apply_q(T._h, ctypes.c_void_p(dC.data_ptr()), n, b, 1)
    assert st == 3  # AVS_ERR_STATE: no matrix reduced yet
    avsvd.Blockbidiag_(T)
    assert lib().avs_apply_q(T._h, ctypes.c_void_p(dC.data_ptr()), n, b, 1) == 3
    assert lib().avs_apply_q(T._h, ctypes.c_void_p(dC.data_ptr()), n, 7, 1) == -4
