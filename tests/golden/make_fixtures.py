"""Generate the committed golden fixtures (small, seeded) with LAPACK.

Run from the repository root:  python tests/golden/make_fixtures.py
Inputs are seeded Philox4x32-10 + Box-Muller matrices (the oracle's C
generator); every expected output comes from scipy's bundled OpenBLAS LAPACK
(dgeqrf, dormqr, dgesdd), never from the reference or the oracle:

  kernels_b{32,64}.npz  GEQRT / TSQRT factors and tau (dgeqrf), UNMQR and
                        TSMQR products Q^T C, Q C, C Q, C Q^T (dormqr on the
                        same factors; TSQRT's Q is the Q of the stacked
                        [triu(A1); A2] factorisation)
  sigma_small.npz       singular values (dgesdd) of the seeded inputs used by
                        the band-reduction tests (n in 16..1024)
"""
import os
import sys

import numpy as np
import scipy.linalg as sl
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

BASE = 20261203


def mat(m, n, seed):
    return oracle.philox_normal(seed, m * n).reshape((m, n), order="F")


def q_apply(factors, tau, c, trans, side="L"):
    out, _, info = lapack.dormqr(side, "T" if trans else "N", factors, tau, np.asfortranarray(c),
                                 max(1, 64 * max(c.shape)))
    assert info == 0
    return out


def kernel_fixture(b):
    a = mat(b, b, BASE + 100 + b)
    qr, tau, _, info = lapack.dgeqrf(a)
    assert info == 0
    w = 5 if b <= 32 else 2  # column tiles in the panel products (keeps files small)
    c = mat(b, w * b, BASE + 200 + b)
    cr = mat(b, b, BASE + 250 + b)
    a1 = mat(b, b, BASE + 300 + b)
    a2 = mat(b, b, BASE + 400 + b)
    st = np.vstack([np.triu(a1), a2])
    sqr, stau, _, info = lapack.dgeqrf(st)
    assert info == 0
    c12 = mat(2 * b, w * b, BASE + 500 + b)
    r12 = mat(b, 2 * b, BASE + 600 + b)
    return dict(
        a=a, geqrt_factors=qr, geqrt_tau=tau,
        c=c, unmqr_QtC=q_apply(qr, tau, c, True), unmqr_QC=q_apply(qr, tau, c, False),
        cr=cr, unmqr_CQ=q_apply(qr, tau, cr, False, "R"), unmqr_CQt=q_apply(qr, tau, cr, True, "R"),
        a1=a1, a2=a2, tsqrt_factors=sqr, tsqrt_tau=stau,
        c12=c12, tsmqr_QtC=q_apply(sqr, stau, c12, True), tsmqr_QC=q_apply(sqr, stau, c12, False),
        r12=r12, tsmqr_CQ=q_apply(sqr, stau, r12, False, "R"), tsmqr_CQt=q_apply(sqr, stau, r12, True, "R"),
    )


def main():
    for b in (32, 64):
        np.savez_compressed(os.path.join(HERE, "kernels_b%d.npz" % b), **kernel_fixture(b))
    sig = {}
    for n in (16, 24, 64, 96, 128, 256, 512, 1024):
        a = oracle.gaussian_matrix(n, BASE + n)
        sig["gauss_n%d" % n] = sl.svd(a, compute_uv=False, lapack_driver="gesdd")
    # the CPU plain-path workload: n=1024, seed index 1
    sig["gauss_n1024_seed1"] = sl.svd(oracle.gaussian_matrix(1024, BASE + 1), compute_uv=False,
                                  lapack_driver="gesdd")
    np.savez_compressed(os.path.join(HERE, "sigma_small.npz"), **sig)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
