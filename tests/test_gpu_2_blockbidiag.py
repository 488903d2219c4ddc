"""Blockbidiag! / tiled_QR! / svdvals on the GPU vs the oracle and LAPACK.

Mirrors KAbasedSVD/test/test_tiledalgo.jl (sigma(band) = sigma(A), R of the
tiled QR = LAPACK R up to signs) and CUDAbasedSVD/test.jl (svdvals through
band + bulge chasing + bdsdc), for both executors (stream-ordered phase
launches and host-enqueued per-task launches), both trees (the reference's
flat tree and the hierarchical one), Float64 and Float32.  Bars: sigma error vs FP64 LAPACK
<= 10*n*eps of the working precision; band structure exact.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import avsvd
import oracle as O

pytestmark = pytest.mark.gpu
BASE = 20261203


def bar(dt, n):
    return 10 * n * np.finfo(dt).eps


def band_of(A, b):
    n = A.shape[0]
    out = np.zeros_like(A)
    for d in range(b + 1):
        idx = np.arange(n - d)
        out[idx, idx + d] = A[idx, idx + d]
    return out


CASES = [
    # (n, b, dtype, sched, tree, domain)
    (64, 32, np.float64, avsvd.SCHED_HOST, avsvd.TREE_FLAT, 0),
    (256, 64, np.float64, avsvd.SCHED_HOST, avsvd.TREE_HIER, 2),
    (256, 32, np.float32, avsvd.SCHED_HOST, avsvd.TREE_HIER, 2),
    (64, 32, np.float64, avsvd.SCHED_STREAMS, avsvd.TREE_FLAT, 0),
    (256, 32, np.float64, avsvd.SCHED_STREAMS, avsvd.TREE_FLAT, 0),
    (256, 64, np.float64, avsvd.SCHED_STREAMS, avsvd.TREE_HIER, 2),
    (512, 64, np.float32, avsvd.SCHED_STREAMS, avsvd.TREE_FLAT, 0),
    (512, 128, np.float64, avsvd.SCHED_STREAMS, avsvd.TREE_HIER, 2),
    (1024, 128, np.float64, avsvd.SCHED_STREAMS, avsvd.TREE_HIER, 4),
    (1024, 64, np.float32, avsvd.SCHED_STREAMS, avsvd.TREE_HIER, 3),
]

PERSISTENT_OFF = ("persistent dependency-counting executor disabled: its launch did not complete on the "
                  "device (twice) and the cause is not identified")


@pytest.mark.skip(reason=PERSISTENT_OFF)
@pytest.mark.parametrize("n,b,dt,tree,domain", [(256, 32, np.float64, avsvd.TREE_FLAT, 0),
                                                (1024, 128, np.float64, avsvd.TREE_HIER, 2)])
def test_blockbidiag_sigma_persistent(gpu, golden_dir, n, b, dt, tree, domain):
    test_blockbidiag_sigma(gpu, golden_dir, n, b, dt, avsvd.SCHED_PERSISTENT, tree, domain)


def test_persistent_executor_reports_disabled(gpu):
    A = np.asfortranarray(np.eye(64))
    T = avsvd.TiledMatrix(A, 32, 32, sched=avsvd.SCHED_PERSISTENT)
    with pytest.raises(avsvd.AvsError):
        avsvd.Blockbidiag_(T)


@pytest.mark.parametrize("n,b,dt,sched,tree,domain", CASES)
def test_blockbidiag_sigma(gpu, golden_dir, n, b, dt, sched, tree, domain):
    a64 = O.gaussian_matrix(n, BASE + n)
    A = np.asfortranarray(a64.astype(dt))
    print("case n=%d b=%d %s sched=%d tree=%d domain=%d" % (n, b, np.dtype(dt).name, sched, tree, domain),
          flush=True)
    T = avsvd.TiledMatrix(A, b, b, sched=sched, tree=tree, domain=domain, wait_ms=10000)
    avsvd.Blockbidiag_(T)
    out = T.A.astype(np.float64)
    # exact zeros outside the band (the reference zeroes V tiles)
    np.testing.assert_array_equal(out, band_of(out, b))
    ref = np.load(os.path.join(golden_dir, "sigma_small.npz"))["gauss_n%d" % n] if n <= 1024 else sl.svdvals(a64)
    err = O.sigma_err(sl.svdvals(out), ref)
    print("sigma err vs FP64 LAPACK: %.3e (bar %.3e)" % (err, bar(dt, n)), flush=True)
    assert err <= bar(dt, n)
    # the band the device hands to the host stage equals the band of A
    bnd = T.band()
    for d in range(b + 1):
        np.testing.assert_array_equal(bnd[b - d, d:], np.diagonal(out, d))


@pytest.mark.parametrize("n,b", [(256, 32), (512, 128)])
def test_flat_tree_band_matches_oracle(gpu, n, b):
    """The reference schedule on the device gives the reference's band (same
    reflectors; only the compact-WY arithmetic order differs), and both
    executors give the same bits (same tile kernels in the same per-tile order)."""
    a = O.gaussian_matrix(n, BASE + 3 * n)
    ref = O.blockbidiag(a, b)
    outs = []
    for sched in (avsvd.SCHED_HOST, avsvd.SCHED_STREAMS):
        print("case n=%d b=%d sched=%d" % (n, b, sched), flush=True)
        A = np.asfortranarray(a.copy())
        T = avsvd.TiledMatrix(A, b, b, sched=sched, tree=avsvd.TREE_FLAT, wait_ms=10000)
        avsvd.Blockbidiag_(T)
        np.testing.assert_allclose(T.A, ref, rtol=0, atol=1e-11 * np.abs(a).max() * np.sqrt(n))
        outs.append(T.A.copy())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n,b,dt,tree", [(1024, 64, np.float64, avsvd.TREE_HIER), (768, 128, np.float64, avsvd.TREE_HIER),
                                         (512, 32, np.float32, avsvd.TREE_FLAT)])
def test_panel_stream_overlap_is_bitwise_neutral(gpu, n, b, dt, tree):
    """Panels running ahead on their own stream change no bit of the result:
    every tile sees the same operations in the same order as on one stream."""
    a = np.asfortranarray(O.gaussian_matrix(n, BASE + 13 * n).astype(dt))
    outs = []
    for single in (True, False):
        print("case n=%d b=%d %s single_stream=%s" % (n, b, np.dtype(dt).name, single), flush=True)
        A = np.asfortranarray(a.copy())
        T = avsvd.TiledMatrix(A, b, b, sched=avsvd.SCHED_STREAMS, tree=tree, single_stream=single)
        avsvd.Blockbidiag_(T)
        outs.append(T.A.copy())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_concurrent_handles_share_panel_streams_bitwise(gpu):
    """Six handles reduced at once (every run enqueued before any is waited
    for), so their twelve streams share the runtime's hardware queues (or, with
    AVSVD_PANEL_QUEUES set, six runs share the pool's panel queues): every band
    equals the same matrix reduced by its handle alone."""
    import ctypes
    from avsvd._lib import Opts, check, dptr, lib
    L = lib()
    n, b, k = 512, 64, 6
    mats = [np.asfortranarray(O.gaussian_matrix(n, BASE + 31 + i)) for i in range(k)]
    o = Opts()
    L.avs_default_opts(ctypes.byref(o))
    hs = []
    try:
        for _ in range(k):
            h = ctypes.c_void_p()
            check(L.avs_create(ctypes.byref(h), avsvd.AVS_F64, n, b, ctypes.byref(o)))
            hs.append(h)

        def bands(together):
            print("case n=%d b=%d handles=%d together=%s" % (n, b, k, together), flush=True)
            for h, a in zip(hs, mats):
                check(L.avs_set_matrix_host(h, a.ctypes.data_as(ctypes.c_void_p), n))
                check(L.avs_synchronize(h))
            for h in hs:
                check(L.avs_blockbidiag(h, 0))
                if not together:
                    check(L.avs_synchronize(h))
            out = []
            for h in hs:
                band = np.zeros((b + 1, n), order="F")
                check(L.avs_get_band(h, dptr(band), b + 1))
                out.append(band)
            return out

        alone = bands(False)
        together = bands(True)
    finally:
        for h in hs:
            L.avs_destroy(h)
    for x, y in zip(alone, together):
        np.testing.assert_array_equal(x, y)
    s = np.linalg.svd(mats[0], compute_uv=False)
    full = np.zeros((n, n))
    for d in range(b + 1):
        idx = np.arange(n - d)
        full[idx, idx + d] = alone[0][b - d, idx + d]
    assert np.max(np.abs(np.linalg.svd(full, compute_uv=False) - s)) <= bar(np.float64, n) * s[0]


def test_pooled_panel_queues_bitwise(gpu):
    """AVSVD_PANEL_QUEUES=2 (read once per process, so in a child process):
    panel streams from a pool of two dedicated hardware queues, taken per run;
    three handles reduced at once give the bands each gives alone."""
    import subprocess
    import sys
    code = r"""
import ctypes, numpy as np, avsvd, oracle as O
from avsvd._lib import Opts, check, dptr, lib
L = lib(); n, b, k = 512, 64, 3
mats = [np.asfortranarray(O.gaussian_matrix(n, 20261203 + 71 + i)) for i in range(k)]
o = Opts(); L.avs_default_opts(ctypes.byref(o))
hs = []
for _ in range(k):
    h = ctypes.c_void_p(); check(L.avs_create(ctypes.byref(h), avsvd.AVS_F64, n, b, ctypes.byref(o))); hs.append(h)
def bands(together):
    for h, a in zip(hs, mats):
        check(L.avs_set_matrix_host(h, a.ctypes.data_as(ctypes.c_void_p), n)); check(L.avs_synchronize(h))
    for h in hs:
        check(L.avs_blockbidiag(h, 0))
        if not together: check(L.avs_synchronize(h))
    out = []
    for h in hs:
        band = np.zeros((b + 1, n), order="F"); check(L.avs_get_band(h, dptr(band), b + 1)); out.append(band)
    return out
alone, together = bands(False), bands(True)
for h in hs: L.avs_destroy(h)
assert all(np.array_equal(x, y) for x, y in zip(alone, together))
print("pooled panel queues OK", flush=True)
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, AVSVD_PANEL_QUEUES="2",
               PYTHONPATH=os.pathsep.join([os.path.join(root, "avocado-sandwich_amd"), os.path.join(root, "oracle"),
                                           root, os.environ.get("PYTHONPATH", "")]))
    print("case n=512 b=64 handles=3 AVSVD_PANEL_QUEUES=2 (child process)", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "pooled panel queues OK" in r.stdout


def test_phase_timing_switch_per_run(gpu):
    """avs_set_phase_timing turns the per-launch events on or off for the next
    runs; the readout describes the last run only and is refused after a run
    without events; the events change no bit of the band."""
    import ctypes
    from avsvd._lib import Opts, check, dptr, lib
    L = lib()
    n, b = 512, 64
    a = np.asfortranarray(O.gaussian_matrix(n, BASE + 41))
    o = Opts()
    L.avs_default_opts(ctypes.byref(o))
    h = ctypes.c_void_p()
    check(L.avs_create(ctypes.byref(h), avsvd.AVS_F64, n, b, ctypes.byref(o)))
    try:
        pm, um, uf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        pn, un = ctypes.c_int64(), ctypes.c_int64()

        def run(timing):
            print("case n=%d b=%d phase_timing=%d" % (n, b, timing), flush=True)
            check(L.avs_set_phase_timing(h, timing))
            check(L.avs_set_matrix_host(h, a.ctypes.data_as(ctypes.c_void_p), n))
            check(L.avs_blockbidiag(h, 0))
            band = np.zeros((b + 1, n), order="F")
            check(L.avs_get_band(h, dptr(band), b + 1))
            st = L.avs_phase_timing(h, ctypes.byref(pm), ctypes.byref(pn), ctypes.byref(um), ctypes.byref(un),
                                    ctypes.byref(uf))
            return band, st

        b0, st0 = run(0)
        b1, st1 = run(1)
        assert st1 == 0 and pn.value > 0 and un.value > 0 and um.value > 0 and uf.value > 0
        check(L.avs_set_phase_timing(h, 0))
        st_stale = L.avs_phase_timing(h, ctypes.byref(pm), ctypes.byref(pn), ctypes.byref(um), ctypes.byref(un),
                                      ctypes.byref(uf))
        b2, st2 = run(0)
        assert st0 != 0 and st2 != 0
        assert st_stale == 0  # the last run (with events) is still the one described
        np.testing.assert_array_equal(b0, b1)
        np.testing.assert_array_equal(b0, b2)
    finally:
        L.avs_destroy(h)


def test_reserved_cus_are_bitwise_neutral(gpu):
    """Updates on a CU-masked stream (16 CUs left to the panels), all of them
    or only the TSQRT updates (the bench configuration), the merges or both, give the same bits as the default
    placement."""
    n, b = 1024, 128
    a = np.asfortranarray(O.gaussian_matrix(n, BASE + 29))
    outs = []
    for k, modes in ((0, 0), (16, 0), (16, 2), (16, 4), (16, 6)):
        print("case n=%d b=%d reserve_cus=%d modes=%d" % (n, b, k, modes), flush=True)
        T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, sched=avsvd.SCHED_STREAMS, reserve_cus=k,
                              masked_modes=modes)
        avsvd.Blockbidiag_(T)
        outs.append(T.A.copy())
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)


@pytest.mark.parametrize("sched,tree", [(avsvd.SCHED_HOST, avsvd.TREE_FLAT),
                                        (avsvd.SCHED_HOST, avsvd.TREE_HIER),
                                        (avsvd.SCHED_STREAMS, avsvd.TREE_FLAT),
                                        (avsvd.SCHED_STREAMS, avsvd.TREE_HIER)])
def test_tiled_qr_R(gpu, sched, tree):
    n, b = 384, 64
    a = O.gaussian_matrix(n, BASE + 9 * n)
    A = np.asfortranarray(a.copy())
    T = avsvd.TiledMatrix(A, b, b, sched=sched, tree=tree, domain=2, wait_ms=10000)
    avsvd.tiled_QR_(T)
    r = sl.qr(a, mode="r")[0]
    out = np.triu(T.A)
    np.testing.assert_allclose(out * np.sign(np.diag(out))[:, None], r * np.sign(np.diag(r))[:, None], atol=1e-10)


@pytest.mark.parametrize("sched", [avsvd.SCHED_HOST, avsvd.SCHED_STREAMS])
def test_kend_partial_matches_oracle(gpu, sched):
    n, b, kend = 256, 32, 3
    a = O.gaussian_matrix(n, BASE + 77)
    ref = O.blockbidiag(a, b, kend=kend)
    A = np.asfortranarray(a.copy())
    T = avsvd.TiledMatrix(A, b, b, sched=sched, tree=avsvd.TREE_FLAT)
    avsvd.Blockbidiag_(T, kend=kend)
    np.testing.assert_allclose(T.A, ref, atol=1e-11 * np.sqrt(n))


def test_n1024_svdvals_end_to_end(gpu, golden_dir):
    """The CPU plain-path workload (n=1024, b=64, FP64, seed index 1) on the device path."""
    n, b = 1024, 64
    a = np.asfortranarray(O.gaussian_matrix(n, BASE + 1))
    ref = np.load(os.path.join(golden_dir, "sigma_small.npz"))["gauss_n1024_seed1"]
    sig = avsvd.svdvals(a, tilesize=b)
    err = O.sigma_err(sig, ref)
    print("n1024 sigma err %.3e" % err, flush=True)
    assert err <= 10 * n * np.finfo(np.float64).eps


def test_n4096_fp64_b128_against_cached_gesdd(gpu, golden_dir):
    """Full-size property checks: ||band||_F = ||A||_F and sigma vs the cached
    FP64 dgesdd vector of the same seeded matrix."""
    n, b = 4096, 128
    a = avsvd.gaussian_matrix(n, BASE)
    fro = np.linalg.norm(a)
    T = avsvd.TiledMatrix(a, b, b, sched=avsvd.SCHED_STREAMS, tree=avsvd.TREE_HIER)
    avsvd.Blockbidiag_(T, sync=False)
    bnd = T.band()
    assert abs(np.linalg.norm(bnd) - fro) <= 1e-12 * fro
    sig = T.svdvals()
    ref = np.load(os.path.join(golden_dir, "sigma_gauss_n4096_s%d.npy" % BASE))
    err = O.sigma_err(sig, ref)
    print("n=4096 sigma err %.3e" % err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_batched_mixed_sizes_with_padding(gpu, dt):
    """Grouped launch over mixed sizes (some not multiples of b): every
    matrix's sigma vs FP64 LAPACK, and each result equals the one-matrix path
    of the same input (same schedule, same arithmetic)."""
    rng = np.random.default_rng(11)
    ns = [64, 40, 128, 100, 256, 32, 96, 160]
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n)).astype(dt) for n in ns]
    bt = avsvd.Batch(ns, dtype=dt, b=32)
    for i, m in enumerate(mats):
        bt.set(i, m)
    sig = bt.svdvals()
    for n, m, s in zip(ns, mats, sig):
        ref = sl.svdvals(m.astype(np.float64))
        err = O.sigma_err(s, ref)
        assert err <= 10 * n * np.finfo(dt).eps, (n, err)
    # band of a multiple-of-b matrix equals the single-matrix host-enqueued flat path
    A = np.asfortranarray(mats[2].copy())
    T = avsvd.TiledMatrix(A, 32, 32, sched=avsvd.SCHED_HOST, tree=avsvd.TREE_FLAT)
    avsvd.Blockbidiag_(T, sync=False)
    np.testing.assert_array_equal(bt.band(2), T.band())


@pytest.mark.parametrize("n,b,dt", [(1024, 128, np.float64), (512, 32, np.float32)])
def test_graph_replays_match_eager(gpu, n, b, dt):
    """The stream-ordered run replayed from its captured HIP graph gives the
    eager run's band bit for bit, on every replay, with eager uploads and
    synchronisations between the replays."""
    a = O.gaussian_matrix(n, BASE + 7 * n + b).astype(dt)
    E = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b)
    avsvd.Blockbidiag_(E, sync=False)
    ref = E.band()
    del E
    T = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, graph=True)
    for rep in range(4):  # capture + replay, then replays
        print("case graph n=%d b=%d rep=%d" % (n, b, rep), flush=True)
        T.A[:] = a
        T.reload()
        avsvd.Blockbidiag_(T, sync=False)
        np.testing.assert_array_equal(T.band(), ref)
        T.synchronize()


@pytest.mark.parametrize("n,b,dt", [(1024, 64, np.float64), (640, 32, np.float32), (96, 32, np.float64)])
def test_svdvals_device_bisection_matches_dbdsdc(gpu, monkeypatch, n, b, dt):
    """The plain path's last step, the bidiagonal's singular values: bisection
    on the device (default from n = 128) and LAPACK dbdsdc on the host agree to
    10*n*eps of FP64, and both match FP64 LAPACK to the working precision's bar."""
    a = O.gaussian_matrix(n, BASE + 11 * n + b)
    outs = []
    for pol in ("", "host"):
        print("case svdvals n=%d b=%d sigma=%s" % (n, b, pol or "device"), flush=True)
        if pol:
            monkeypatch.setenv("AVSVD_BDSDC", pol)
        else:
            monkeypatch.delenv("AVSVD_BDSDC", raising=False)
        outs.append(avsvd.svdvals(a.astype(dt), tilesize=b))
    assert O.sigma_err(outs[0], outs[1]) <= 10 * n * np.finfo(np.float64).eps
    ref = sl.svdvals(a.astype(dt).astype(np.float64))
    for o in outs:
        assert O.sigma_err(o, ref) <= bar(dt, n)


def test_batched_device_stage2_matches_host(gpu, monkeypatch):
    """The batch's host stage on the device (the default for b <= 32): band ->
    bidiagonal by bulge chasing (one wavefront per matrix) and the bidiagonal's
    singular values by bisection (one lane per value).  Each device step gives
    the sigma of its host counterpart (host bulge chasing, LAPACK dbdsdc) to
    10*n*eps of FP64, and all match FP64 LAPACK to that bar."""
    rng = np.random.default_rng(17)
    ns = [1, 2, 31, 32, 33, 64, 97, 200, 256, 300, 1000]
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n)) for n in ns]
    mats[6][:, 3] = 0.0  # a rank-deficient matrix: an exact zero singular value
    outs = []
    for stage2, bdsdc in (("", ""), ("", "host"), ("host", "")):
        print("case batched stage2=%s sigma=%s" % (stage2 or "device", "host" if stage2 or bdsdc else "device"),
              flush=True)
        for var, val in (("AVSVD_STAGE2", stage2), ("AVSVD_BDSDC", bdsdc)):
            if val:
                monkeypatch.setenv(var, val)
            else:
                monkeypatch.delenv(var, raising=False)
        bt = avsvd.Batch(ns, dtype=np.float64, b=32)
        for i, m in enumerate(mats):
            bt.set(i, m)
        outs.append(bt.svdvals())
        del bt
    for k, (n, m) in enumerate(zip(ns, mats)):
        bar = 10 * max(n, 2) * np.finfo(np.float64).eps
        ref = sl.svdvals(m)
        for o in outs:
            assert len(o[k]) == n
            assert O.sigma_err(o[k], ref) <= bar, n
        assert O.sigma_err(outs[0][k], outs[1][k]) <= bar, n
        assert O.sigma_err(outs[1][k], outs[2][k]) <= bar, n


SERVING_HUNG = ("a full -m gpu run went silent in this test (killed after 180 s) with the serving loop's earlier "
                "design (one thread per in-flight reduction, each making its own HIP calls); the cause was not "
                "identified, so this case stays off the GPU; the redesigned loop (one device thread makes every "
                "HIP call) is covered by tests/test_gpu_5_parity.py::test_serving_*")


@pytest.mark.skip(reason=SERVING_HUNG)
def test_serving_results_same_alone_and_in_mix(gpu):
    """Continuous batching: each request's sigma is the same when it is served
    alone and when it overlaps with other requests on the device (the task DAG
    fixes the order of every operation on each tile), and matches LAPACK."""
    rng = np.random.default_rng(5)
    sizes = [(300, np.float64), (130, np.float32), (512, np.float64), (96, np.float32), (700, np.float32)]
    mats = [rng.standard_normal((n, n)).astype(dt) for n, dt in sizes]
    srv = avsvd.Server(hbm_budget_gb=8, max_inflight=3, host_threads=4)
    alone = []
    for m in mats:
        alone.append(srv.wait(srv.submit(m))[0])
    ids = [srv.submit(m) for m in mats]
    mixed = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 2 * len(mats)
    for (n, dt), m, a, x in zip(sizes, mats, alone, mixed):
        np.testing.assert_array_equal(a, x)
        err = O.sigma_err(x, sl.svdvals(m.astype(np.float64)))
        assert err <= 10 * n * np.finfo(dt).eps, (n, dt, err)
