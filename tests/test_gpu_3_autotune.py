"""The autotuner on the device: it times every candidate, and the tile size it
picks gives the right singular values (bar: 10*n*eps vs FP64 LAPACK)."""
import numpy as np
import pytest
import scipy.linalg as sl

import avsvd
import oracle as O
from avsvd import autotune as at

pytestmark = pytest.mark.gpu


def test_tune_plain_and_use_choice(gpu):
    n = 512
    print("case tune_plain n=%d" % n, flush=True)
    for form in ("band", "svdvals"):
        e = at.tune_plain(n, np.float64, form=form, tiles=(32, 64, 128), domains=(0, 2), reps=1)
        assert sorted((t[0], t[1]) for t in e["trials"]) == [(32, 0), (32, 2), (64, 0), (64, 2), (128, 0), (128, 2)]
        assert all(t[2] > 0 for t in e["trials"])
        assert e["ms"] == min(t[2] for t in e["trials"]) and e["b"] in at.TILES
    arch = at.device_arch()
    assert arch and arch.startswith("gfx")
    t = at.TuneTable()
    t.record(arch, np.float64, "svdvals", n, e)
    b, dom = at.choose_tile(n, np.float64, "svdvals", table=t, arch=arch)
    assert (b, dom) == (e["b"], e["domain"])
    a = O.gaussian_matrix(n, 20261203 + 71)
    sig = avsvd.TiledMatrix(np.asfortranarray(a.copy()), b, b, domain=dom).svdvals()
    assert O.sigma_err(sig, sl.svdvals(a)) <= 10 * n * np.finfo(np.float64).eps


def test_tune_batch_and_serving_table(gpu, tmp_path):
    ns = [64, 128, 256, 96, 200, 64]
    print("case tune_batch ns=%s" % ns, flush=True)
    e = at.tune_batch(ns, np.float32, tiles=(32, 64), reps=1)
    assert [t[0] for t in e["trials"]] == [32, 64] and e["b"] in (32, 64)
    # a served request takes the table's tile size
    arch = at.device_arch()
    t = at.TuneTable()
    t.record(arch, np.float32, "svdvals", 256, {"b": 64, "domain": 0, "ms": 1.0, "trials": []})
    p = str(tmp_path / "tune.json")
    t.save(p)
    try:
        m = np.random.default_rng(3).standard_normal((200, 200)).astype(np.float32)
        # the server loads the table into the library
        srv = avsvd.Server(hbm_budget_gb=4, max_inflight=2, host_threads=2, tune_table=p)
        from avsvd.server import served_tile
        assert served_tile(np.float32, 200) == 64 and served_tile(np.float64, 200) == 32
        sig = srv.wait(srv.submit(m))[0]
        srv.close()
        assert O.sigma_err(sig, sl.svdvals(m.astype(np.float64))) <= 10 * 200 * np.finfo(np.float32).eps
    finally:
        avsvd.lib().avs_tune_clear()
