"""Band -> bidiagonal on the GPU for one large matrix (the pipelined chase,
avs_bidiagonalize_band_device) against the host chase of the same band
(avs_bidiagonalize_band_chase, whose singular values are pinned against
LAPACK in test_abi.py) and against FP64 LAPACK svdvals of the dense band.

The reference runs this stage on the host (KAbasedSVD/src/bulgechasing.jl:33-42)
and has no device counterpart.  Bars: singular values of (d, e) within 10*n*eps
of FP64 LAPACK and of the host chase's.  d and e themselves are not a parity
target: the device applies the same reflectors in the same order but rounds
the dot products in another order, and single entries of d and e can be far
more sensitive than the singular values (n=2048, bw=2: a 1e-16 relative
perturbation of the band moves an entry of d by 3e-8 on the host chase itself).
"""
import numpy as np
import pytest
import scipy.linalg as sl

import avsvd
import oracle as O

pytestmark = pytest.mark.gpu
BASE = 20261211


def banded(n, bw, seed, grade=None):
    a = O.gaussian_matrix(max(n, 2), seed)[:n, :n]
    if grade is not None:
        a = a * grade[:, None]
    return np.triu(np.tril(a, bw))


@pytest.mark.parametrize("n,bw", [(1, 4), (2, 4), (5, 4), (7, 3), (64, 8), (100, 1), (257, 31), (300, 32),
                                  (700, 16), (1000, 64), (600, 128), (1500, 128), (2048, 2)])
def test_device_chase_matches_host_chase(gpu, n, bw):
    print("case device chase n=%d bw=%d" % (n, bw), flush=True)
    band = banded(n, bw, BASE + 7 * n + bw)
    st = avsvd.tiled._dense_to_band(band, bw)
    d, e = avsvd.tiled.bidiagonalize_band_device(st, n, bw)
    dh, eh = avsvd.tiled.bidiagonalize_band_chase(st, n, bw)
    assert np.isfinite(d).all() and np.isfinite(e).all()
    sig = avsvd.bdsdc_values(d, e)
    ref = sl.svdvals(band) if n > 1 else np.abs(np.diag(band))
    assert O.sigma_err(sig, ref) < 10 * max(n, 2) * np.finfo(float).eps
    assert O.sigma_err(sig, avsvd.bdsdc_values(dh, eh)) < 10 * max(n, 2) * np.finfo(float).eps


def test_device_chase_graded_band(gpu):
    """Entries spanning 1e-300 .. 1: the chased bulges decay to subnormal sizes
    without Inf or NaN (divisions, not reciprocals, as on the host)."""
    n, bw = 600, 16
    print("case device chase graded n=%d bw=%d" % (n, bw), flush=True)
    band = banded(n, bw, 20261237, grade=np.logspace(0, -300, n))
    d, e = avsvd.tiled.bidiagonalize_band_device(avsvd.tiled._dense_to_band(band, bw), n, bw)
    assert np.isfinite(d).all() and np.isfinite(e).all()
    assert O.sigma_err(avsvd.bdsdc_values(d, e), sl.svdvals(band)) < 10 * n * np.finfo(float).eps


@pytest.mark.parametrize("n,b,dt", [(1024, 64, np.float64), (640, 32, np.float32), (512, 128, np.float64)])
def test_svdvals_device_chase_matches_host_chase(gpu, monkeypatch, n, b, dt):
    """svdvals with the band -> bidiagonal stage on the device (from the
    handle's band store) and on the host give the same singular values to
    10*n*eps of FP64, and both match FP64 LAPACK to the working precision's bar."""
    a = O.gaussian_matrix(n, BASE + 13 * n + b)
    outs = []
    for pol in ("device", "host"):
        print("case svdvals n=%d b=%d stage2=%s" % (n, b, pol), flush=True)
        monkeypatch.setenv("AVSVD_STAGE2", pol)
        outs.append(avsvd.svdvals(a.astype(dt), tilesize=b))
    assert O.sigma_err(outs[0], outs[1]) <= 10 * n * np.finfo(np.float64).eps
    ref = sl.svdvals(a.astype(dt).astype(np.float64))
    for o in outs:
        assert O.sigma_err(o, ref) <= 10 * n * np.finfo(dt).eps


def test_chase_variants_child_process(gpu, tmp_path):
    """AVSVD_CHASE_LDS=0 (each step's right-phase rows through HBM instead of
    LDS) and AVSVD_BATCH_CHASE=wave (one wavefront per batched matrix instead
    of a pipelined workgroup) are read once per process, so they run in a
    child process.  The single-matrix chase keeps sigma (its reductions take a
    different order there); the batched chase gives the same bits."""
    import hashlib
    import os
    import subprocess
    import sys
    code = r"""
import hashlib, numpy as np, avsvd, oracle as O
n, bw = 1500, 128
a = np.triu(np.tril(O.gaussian_matrix(n, 20261211 + 3), bw))
d, e = avsvd.tiled.bidiagonalize_band_device(avsvd.tiled._dense_to_band(a, bw), n, bw)
np.save(OUT + "_sig.npy", avsvd.bdsdc_values(d, e))
ns = [33, 200, 300, 1000]
rng = np.random.default_rng(29)
bt = avsvd.Batch(ns, dtype=np.float64, b=32)
for i, k in enumerate(ns):
    bt.set(i, rng.standard_normal((k, k)) / np.sqrt(k))
print("batch", hashlib.sha256(b"".join(s.tobytes() for s in bt.svdvals())).hexdigest(), flush=True)
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = str(tmp_path / "chase_variants")
    env = dict(os.environ, AVSVD_CHASE_LDS="0", AVSVD_BATCH_CHASE="wave",
               PYTHONPATH=os.pathsep.join([os.path.join(root, "avocado-sandwich_amd"), os.path.join(root, "oracle"),
                                           root, os.environ.get("PYTHONPATH", "")]))
    print("case chase variants n=1500 bw=128, batch of 4 (child process)", flush=True)
    r = subprocess.run([sys.executable, "-c", "OUT = %r\n" % out + code], env=env, capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    child_batch = [x for x in r.stdout.splitlines() if x.startswith("batch ")][0].split()[1]
    # the same in this process with the defaults
    n, bw = 1500, 128
    a = np.triu(np.tril(O.gaussian_matrix(n, 20261211 + 3), bw))
    d, e = avsvd.tiled.bidiagonalize_band_device(avsvd.tiled._dense_to_band(a, bw), n, bw)
    sig = avsvd.bdsdc_values(d, e)
    assert O.sigma_err(sig, np.load(out + "_sig.npy")) < 10 * n * np.finfo(float).eps
    assert O.sigma_err(sig, sl.svdvals(a)) < 10 * n * np.finfo(float).eps
    ns = [33, 200, 300, 1000]
    rng = np.random.default_rng(29)
    bt = avsvd.Batch(ns, dtype=np.float64, b=32)
    for i, k in enumerate(ns):
        bt.set(i, rng.standard_normal((k, k)) / np.sqrt(k))
    here = hashlib.sha256(b"".join(s.tobytes() for s in bt.svdvals())).hexdigest()
    assert here == child_batch
