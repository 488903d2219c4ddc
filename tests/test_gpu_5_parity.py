"""Parity at the configurations the band reduction is quoted on, and for the
paths that only run there.

- FP32 updates as three-term split-BF16 products (update_mode SPLIT_BF16):
  one Blockbidiag! step (every tile operator in both orientations) against
  the oracle, and whole reductions against FP64 LAPACK.
- Full sizes (n = 8192 and 16384, FP64 b=128, FP32 b=64) against the cached
  FP64 dgesdd vectors of the same seeded Gaussian matrices, through the
  default svdvals path (device bulge chasing from n = 16384, device bisection).
- Prescribed spectra A = U diag(s) V^T (exact sigma = s) at n = 4096 and, for
  the precision form, at n = 16384 with cond 1e2, 1e4, 1e6, FP32 native and
  split-BF16.
- Handle state: svdvals after a partial Blockbidiag!, after tiled_QR!, and a
  second Blockbidiag! on a reduced matrix.
Reference tests mirrored: KAbasedSVD/test/test_tiledalgo.jl:11-69 (sigma of
the band vs LAPACK), KAbasedSVD/benchmark/runbenchmarks.jl:15-34 (the sizes
and element types benchmarked).
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg as sl

import avsvd
import oracle as O
from avsvd import inputs
from avsvd._lib import AvsError, check, dptr, lib
from avsvd.suites import Reducer

pytestmark = pytest.mark.gpu
BASE = 20261203
F32_BAR = 1e-5   # FP32 storage forms: max |sigma_i - ref_i| <= 1e-5 sigma_max
F64_BAR = 1e-12  # FP64 form at full size


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("b", [32, 64, 128])
@pytest.mark.parametrize("sched", [avsvd.SCHED_HOST, avsvd.SCHED_STREAMS])
def test_split_bf16_one_step_matches_oracle(gpu, b, sched):
    """kend = 1: GEQRT, UNMQR, the TSQRT chain and TSMQR on tile column 0,
    then the same on the adjoint of tile row 0 (the LQ twins), with the
    trailing updates as split-BF16 products.  Against the FP64 oracle of the
    FP32-rounded input; the native FP32 updates of the same step are checked
    to their tighter bar in the same run."""
    n = 4 * b
    a = O.gaussian_matrix(n, BASE + 31 * b).astype(np.float32)
    ref = O.blockbidiag(a.astype(np.float64), b, kend=1)
    scale = float(np.abs(ref).max())
    errs = {}
    for mode, label in ((avsvd.UPD_NATIVE, "native"), (avsvd.UPD_SPLIT_BF16, "split")):
        _say("case one-step b=%d sched=%d %s" % (b, sched, label))
        T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, sched=sched, tree=avsvd.TREE_FLAT,
                              update_mode=mode)
        avsvd.Blockbidiag_(T, kend=1)
        errs[label] = float(np.abs(T.A.astype(np.float64) - ref).max()) / scale
        del T
    _say("b=%d native %.3e split %.3e" % (b, errs["native"], errs["split"]))
    # native FP32: a few roundings of O(b) accumulated products; split-BF16:
    # each product carries ~2^-16 relative error (the dropped lo*lo term and
    # the unrepresented tail of the split)
    assert errs["native"] <= 20 * b * np.finfo(np.float32).eps
    assert errs["split"] <= 4 * b * 2.0 ** -16
    assert errs["split"] >= errs["native"] * 0.5  # the split path really ran


@pytest.mark.parametrize("b,tree", [(32, avsvd.TREE_HIER), (64, avsvd.TREE_FLAT), (128, avsvd.TREE_HIER)])
def test_split_bf16_blockbidiag_sigma_vs_gesdd(gpu, b, tree):
    n = 1024
    a = O.gaussian_matrix(n, BASE + 5 * b).astype(np.float32)
    _say("case split sigma n=%d b=%d tree=%d" % (n, b, tree))
    T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, tree=tree, update_mode=avsvd.UPD_SPLIT_BF16)
    sig = T.svdvals()
    err = O.sigma_err(sig, sl.svdvals(a.astype(np.float64)))
    _say("split n=%d b=%d sigma err %.3e" % (n, b, err))
    assert err <= F32_BAR


def test_split_bf16_batched(gpu):
    rng = np.random.default_rng(23)
    ns = [64, 100, 256, 512, 32, 200]
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32) for n in ns]
    _say("case batched split")
    bt = avsvd.Batch(ns, dtype=np.float32, b=32, update_mode=avsvd.UPD_SPLIT_BF16)
    for i, m in enumerate(mats):
        bt.set(i, m)
    sig = bt.svdvals()
    for n, m, s in zip(ns, mats, sig):
        assert O.sigma_err(s, sl.svdvals(m.astype(np.float64))) <= F32_BAR, n


def _golden(golden_dir, n):
    return np.load(os.path.join(golden_dir, "sigma_gauss_n%d_s%d.npy" % (n, BASE)))


@pytest.mark.parametrize("n,b,dt", [(8192, 128, np.float64), (16384, 128, np.float64),
                                    (8192, 64, np.float32), (16384, 64, np.float32)])
def test_full_size_gaussian_vs_cached_gesdd(gpu, golden_dir, n, b, dt):
    """The plain GPU configurations at size: Blockbidiag! then the default
    svdvals path against FP64 dgesdd of the same seeded matrix; the band keeps
    ||A||_F exactly up to rounding."""
    import torch
    _say("case full-size n=%d b=%d %s" % (n, b, np.dtype(dt).name))
    a = avsvd.gaussian_matrix(n, BASE, dtype=dt)
    fro = float(np.linalg.norm(a.astype(np.float64)))
    t = torch.from_numpy(a.T).to(gpu)
    torch.cuda.synchronize()
    del a
    R = Reducer(n, b, "f64" if dt == np.float64 else "f32")
    R.reduce(t.data_ptr(), reps=1)
    del t
    band = np.zeros((b + 1, n), order="F")
    check(lib().avs_get_band(R.h, dptr(band), b + 1))
    fro_err = abs(float(np.linalg.norm(band)) - fro) / fro
    sig, secs = R.svdvals()
    R.close()
    err = O.sigma_err(sig, _golden(golden_dir, n))
    _say("n=%d b=%d %s fro %.2e sigma err %.3e (svdvals %.2f s)" % (n, b, np.dtype(dt).name, fro_err, err, secs))
    if dt == np.float64:
        assert fro_err <= 1e-13
        assert err <= F64_BAR
    else:
        assert fro_err <= 1e-5
        assert err <= F32_BAR


@pytest.mark.parametrize("b,dt", [(64, np.float64), (128, np.float64), (64, np.float32)])
def test_geometric_spectrum_n4096_vs_exact(gpu, b, dt):
    """Input (ii) of the plain configurations: A = U diag(s) V^T with Haar U,
    V and s from 1 down to 1e-6; the exact singular values are s."""
    n = 4096
    _say("case geometric n=%d b=%d %s" % (n, b, np.dtype(dt).name))
    import torch
    t, s = inputs.geometric_matrix_device(n, BASE + 3, 1e6, dtype=dt)
    # the handle's stream does not wait for torch's: the input must be complete
    torch.cuda.synchronize()
    R = Reducer(n, b, "f64" if dt == np.float64 else "f32")
    R.reduce(t.data_ptr(), reps=1)
    del t
    sig, _ = R.svdvals()
    R.close()
    err = O.sigma_err(sig, s)
    _say("geometric n=%d b=%d err %.3e" % (n, b, err))
    assert err <= (F64_BAR if dt == np.float64 else F32_BAR)


@pytest.fixture(scope="module")
def haar_pair(gpu):
    n = 16384
    _say("case haar pair n=%d" % n)
    u = inputs._haar_device(n, BASE + 101, "cuda:0")
    v = inputs._haar_device(n, BASE + 101 + 7919, "cuda:0")
    yield u, v
    del u, v
    import torch
    torch.cuda.empty_cache()


@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_precision_form_n16384_f32(gpu, haar_pair, cond):
    """The precision configuration: n=16384 FP32 storage, geometric spectra,
    the trailing updates in native FP32 and as split-BF16 products; every
    singular value within 1e-5 sigma_max of the exact s."""
    import torch
    n, b = 16384, 128
    u, v = haar_pair
    s = inputs.geometric_sigma(n, cond)
    st = torch.from_numpy(s).to(u.device)
    at = ((u * st.unsqueeze(0)) @ v.t().contiguous()).t().contiguous().to(torch.float32)
    torch.cuda.synchronize()
    for mode, label in ((avsvd.UPD_NATIVE, "native"), (avsvd.UPD_SPLIT_BF16, "split")):
        _say("case precision cond=%g %s" % (cond, label))
        R = Reducer(n, b, "f32", update_mode=mode)
        R.reduce(at.data_ptr(), reps=1)
        sig, _ = R.svdvals()
        R.close()
        err = O.sigma_err(sig, s)
        _say("precision cond=%g %s err %.3e" % (cond, label, err))
        assert err <= F32_BAR
    del at


def test_svdvals_after_partial_blockbidiag(gpu):
    """A partial Blockbidiag! leaves an orthogonal transform of the matrix:
    svdvals finishes it and returns the matrix's singular values."""
    n, b = 512, 64
    a = O.gaussian_matrix(n, BASE + 211)
    ref = sl.svdvals(a)
    for sched in (avsvd.SCHED_STREAMS, avsvd.SCHED_HOST):
        _say("case partial then svdvals sched=%d" % sched)
        T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, sched=sched)
        avsvd.Blockbidiag_(T, kend=3)
        assert O.sigma_err(T.svdvals(), ref) <= 10 * n * np.finfo(np.float64).eps


def test_blockbidiag_twice_keeps_sigma(gpu):
    """Blockbidiag! on an already reduced matrix runs on the band with its
    explicit zeros, as the reference's in-place matrix holds it."""
    n, b = 512, 32
    a = O.gaussian_matrix(n, BASE + 212)
    _say("case blockbidiag twice")
    T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b)
    avsvd.Blockbidiag_(T)
    first = T.A.copy()
    avsvd.Blockbidiag_(T)
    second = T.A.copy()
    ref = sl.svdvals(a)
    assert O.sigma_err(sl.svdvals(first), ref) <= 10 * n * np.finfo(np.float64).eps
    assert O.sigma_err(sl.svdvals(second), ref) <= 10 * n * np.finfo(np.float64).eps
    # the second result is a band too: exact zeros outside it
    assert np.array_equal(np.tril(second, -1), np.zeros_like(second))
    assert np.array_equal(np.triu(second, b + 1), np.zeros_like(second))


def test_svdvals_after_tiled_qr_is_refused(gpu):
    n, b = 256, 32
    a = O.gaussian_matrix(n, BASE + 213)
    _say("case tiled_QR then svdvals")
    T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b)
    avsvd.tiled_QR_(T)
    with pytest.raises(AvsError) as ei:
        T.svdvals()
    assert ei.value.status == 3  # AVS_ERR_STATE
    T.A[:] = a  # tiled_QR! wrote R and the reflectors into T.A, as the reference does in place
    T.reload()  # a fresh upload makes the handle usable again
    assert O.sigma_err(T.svdvals(), sl.svdvals(a)) <= 10 * n * np.finfo(np.float64).eps


def test_batched_fp32_large_members(gpu):
    """The batched configuration's largest members (n = 512 and 1024, FP32,
    N(0, 1/n) entries) in one grouped launch with smaller ones."""
    rng = np.random.default_rng(29)
    ns = [1024, 512, 64, 1024, 128, 512, 256]
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32) for n in ns]
    _say("case batched fp32 large members")
    bt = avsvd.Batch(ns, dtype=np.float32, b=32)
    for i, m in enumerate(mats):
        bt.set(i, m)
    sig = bt.svdvals()
    for n, m, s in zip(ns, mats, sig):
        err = O.sigma_err(s, sl.svdvals(m.astype(np.float64)))
        assert err <= F32_BAR, (n, err)


@pytest.mark.parametrize("device_sigma", [False, True])
def test_serving_large_fp64_and_geometric(gpu, device_sigma):
    """Served requests that include an FP64 n=2048 Gaussian matrix and an FP32
    geometric-spectrum matrix (cond 1e4): each result is the same alone and in
    the mix, and within the bar of its reference (LAPACK, or the exact s)."""
    rng = np.random.default_rng(41)
    g64 = rng.standard_normal((2048, 2048))
    geo, s_geo = inputs.geometric_matrix_host(1000, BASE + 17, 1e4, dtype=np.float32)
    reqs = [(g64, sl.svdvals(g64), 10 * 2048 * np.finfo(np.float64).eps),
            (geo, s_geo, F32_BAR),
            (rng.standard_normal((300, 300)).astype(np.float32), None, F32_BAR),
            (rng.standard_normal((130, 130)), None, 10 * 130 * np.finfo(np.float64).eps)]
    srv = avsvd.Server(hbm_budget_gb=16, max_inflight=3, host_threads=4, device_sigma=device_sigma)
    _say("case serving device_sigma=%s alone" % device_sigma)
    alone = [srv.wait(srv.submit(m))[0] for m, _, _ in reqs]
    _say("case serving device_sigma=%s mix" % device_sigma)
    ids = [srv.submit(m) for m, _, _ in reqs]
    mixed = [srv.wait(i)[0] for i in ids]
    srv.close()
    for (m, ref, bar), x, y in zip(reqs, alone, mixed):
        np.testing.assert_array_equal(x, y)
        if ref is None:
            ref = sl.svdvals(m.astype(np.float64))
        assert O.sigma_err(y, ref) <= bar, m.shape


def test_serving_pool_reuse_and_eviction(gpu):
    """Requests of two shapes alternate under a budget that holds only one
    pooled handle at a time: handles are evicted and recreated, and every
    result still matches LAPACK."""
    rng = np.random.default_rng(43)
    mats = [rng.standard_normal((n, n)) for n in (512, 384, 512, 384, 512)]
    # footprint of one n=512 FP64 handle is ~64 MiB + tiles; 0.1 GiB fits one
    srv = avsvd.Server(hbm_budget_gb=0.1, max_inflight=2, host_threads=2, b_f32=64, b_f64=64)
    _say("case serving pool eviction")
    ids = [srv.submit(m) for m in mats]
    outs = [srv.wait(i)[0] for i in ids]
    again = srv.wait(srv.submit(mats[0]))[0]
    srv.close()
    for m, o in zip(mats, outs):
        assert O.sigma_err(o, sl.svdvals(m)) <= 10 * m.shape[0] * np.finfo(np.float64).eps
    np.testing.assert_array_equal(again, outs[0])
