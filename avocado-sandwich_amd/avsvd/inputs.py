"""Seeded synthetic inputs for tests and benchmarks.

- iid N(0, scale^2): Philox4x32-10 + Box-Muller on the host (libavsvd).
- Prescribed spectrum: A = U diag(s) V^T with U, V Haar orthogonal (QR of a
  seeded Gaussian matrix, columns sign-fixed by diag(R)), formed in FP64 on
  the device; s_i = cond^(-(i-1)/(n-1)) decays geometrically from 1 to 1/cond.
  The exact singular values are s (up to the FP64 formation error), so they
  are their own oracle.
Device tensors use torch only as device memory; they hold A column-major
(the row-major storage of A^T), ready for avs_set_matrix_device.
"""
import numpy as np

from .tiled import gaussian_matrix

BASE_SEED = 20261203


def geometric_sigma(n, cond):
    if n == 1:
        return np.ones(1)
    return cond ** (-np.arange(n) / (n - 1.0))


def _haar_device(n, seed, device):
    import torch
    g = torch.from_numpy(gaussian_matrix(n, seed)).to(device)
    q, r = torch.linalg.qr(g)
    d = torch.sign(torch.diagonal(r))
    d[d == 0] = 1.0
    return q * d.unsqueeze(0)


def geometric_matrix_device(n, seed, cond, dtype=np.float64, device="cuda:0"):
    """(device tensor holding A column-major in `dtype`, exact sigma as float64)."""
    import torch
    s = geometric_sigma(n, cond)
    u = _haar_device(n, seed, device)
    v = _haar_device(n, seed + 7919, device)
    st = torch.from_numpy(s).to(device)
    a = (u * st.unsqueeze(0)) @ v.t().contiguous()  # FP64
    del u, v
    at = a.t().contiguous()  # row-major A^T == column-major A
    del a
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    return at.to(tdt), s


def geometric_matrix_host(n, seed, cond, dtype=np.float64, device="cuda:0"):
    """Host (Fortran-ordered) copy of geometric_matrix_device."""
    t, s = geometric_matrix_device(n, seed, cond, dtype, device)
    return np.asfortranarray(t.cpu().numpy().T), s
