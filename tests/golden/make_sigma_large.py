"""Cache FP64 LAPACK dgesdd singular values of the large seeded inputs.

Run from the repository root:  python tests/golden/make_sigma_large.py N [seed]
Writes tests/golden/sigma_gauss_n{N}_s{seed}.npy (descending, float64).
The matrix is the seeded N(0,1) input (Philox4x32-10 + Box-Muller, generated by
the oracle's C generator); scipy's bundled OpenBLAS computes the values.
"""
import os
import sys
import time

import numpy as np
import scipy.linalg

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "..", "oracle"))
import oracle  # noqa: E402

n = int(sys.argv[1])
seed = int(sys.argv[2]) if len(sys.argv) > 2 else 20261203
t0 = time.time()
A = oracle.gaussian_matrix(n, seed)
print("generated", n, time.time() - t0, flush=True)
s = scipy.linalg.svd(A, compute_uv=False, overwrite_a=True, check_finite=False, lapack_driver="gesdd")
out = os.path.join(os.path.dirname(__file__), "sigma_gauss_n%d_s%d.npy" % (n, seed))
np.save(out, s)
print("done", n, time.time() - t0, out, flush=True)
