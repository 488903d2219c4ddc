"""Autotuner table logic and the library's tile-size table (host only)."""
import ctypes

import numpy as np
import pytest

import avsvd
from avsvd import autotune as at
from avsvd._lib import AVS_F32, AVS_F64, lib


def test_bucket_is_next_power_of_two():
    assert [at.bucket(n) for n in (1, 2, 3, 64, 65, 1000, 1024, 1025, 16384)] == \
        [1, 2, 4, 64, 128, 1024, 1024, 2048, 16384]
    with pytest.raises(ValueError):
        at.bucket(0)


def test_table_lookup_round_trip(tmp_path):
    t = at.TuneTable()
    t.record("gfx950", np.float64, "band", 4096, {"b": 64, "domain": 2, "ms": 46.0, "trials": [[64, 2, 46.0]]})
    t.record("gfx950", np.float64, "band", 16384, {"b": 128, "domain": 2, "ms": 478.0, "trials": []})
    t.record("gfx950", np.float32, "band", 16384, {"b": 64, "domain": 0, "ms": 373.0, "trials": []})
    t.record("gfx942", np.float64, "band", 16384, {"b": 32, "domain": 0, "ms": 1.0, "trials": []})
    p = str(tmp_path / "sub" / "tune.json")
    t.save(p)
    u = at.TuneTable.load(p)
    assert u.entries == t.entries
    # smallest bucket >= n, else the largest; arch, dtype and form are separate
    assert u.lookup("gfx950", np.float64, "band", 100)["b"] == 64
    assert u.lookup("gfx950", np.float64, "band", 4096)["b"] == 64
    assert u.lookup("gfx950", np.float64, "band", 4097)["b"] == 128
    assert u.lookup("gfx950", np.float64, "band", 1 << 20)["b"] == 128
    assert u.lookup("gfx950", np.float32, "band", 100)["b"] == 64
    assert u.lookup("gfx942", np.float64, "band", 100)["b"] == 32
    assert u.lookup("gfx950", np.float64, "svdvals", 100) is None
    assert u.buckets("gfx950", np.float64, "band") == [4096, 16384]
    with pytest.raises(ValueError):
        t.record("gfx950", np.float64, "other", 64, {"b": 32})


def test_choose_tile_respects_divisibility_and_defaults():
    t = at.TuneTable()
    t.record("gfx950", np.float64, "band", 2048,
             {"b": 128, "domain": 4, "ms": 1.0, "trials": [[128, 4, 1.0], [32, 0, 3.0], [64, 0, 2.0]]})
    assert at.choose_tile(2048, np.float64, "band", table=t, arch="gfx950") == (128, 4)
    # 1088 = 17 * 64: 128 does not divide it, the next fastest measured tile that does is 64
    assert at.choose_tile(1088, np.float64, "band", table=t, arch="gfx950") == (64, 0)
    # 1056 = 33 * 32
    assert at.choose_tile(1056, np.float64, "band", table=t, arch="gfx950") == (32, 0)
    # no entry: the defaults, reduced to a tile that divides n
    e = at.TuneTable()
    assert at.choose_tile(16384, np.float64, "band", table=e, arch="gfx950") == (128, 0)
    assert at.choose_tile(16384, np.float32, "band", table=e, arch="gfx950") == (64, 0)
    assert at.choose_tile(1024, np.float64, "svdvals", table=e, arch="gfx950") == (32, 0)
    assert at.choose_tile(4160, np.float64, "svdvals", table=e, arch="gfx950") == (64, 0)
    assert at.choose_tile(4128, np.float64, "svdvals", table=e, arch="gfx950") == (32, 0)
    assert at.choose_tile(1000, np.float32, "batch", table=e, arch="gfx950") == (32, 0)
    with pytest.raises(avsvd.AvsError):
        at.choose_tile(1000, np.float64, "band", table=e, arch="gfx950")


def _lookup(code, n):
    b = ctypes.c_int32()
    assert lib().avs_tune_lookup(code, n, ctypes.byref(b)) == 0
    return b.value


def test_library_tile_table():
    L = lib()
    try:
        assert L.avs_tune_clear() == 0
        assert (_lookup(AVS_F64, 1000), _lookup(AVS_F64, 2048), _lookup(AVS_F64, 2049)) == (32, 32, 64)
        assert L.avs_tune_set(AVS_F64, 4096, 128) == 0
        assert L.avs_tune_set(AVS_F64, 1024, 32) == 0
        assert L.avs_tune_set(AVS_F64, 2048, 64) == 0
        assert [_lookup(AVS_F64, n) for n in (1, 1024, 1025, 2048, 4000, 4096, 100000)] == \
            [32, 32, 64, 64, 128, 128, 128]
        assert L.avs_tune_set(AVS_F64, 2048, 128) == 0  # replaces the entry
        assert _lookup(AVS_F64, 1500) == 128
        assert _lookup(AVS_F32, 4096) == 64  # other dtype: defaults
        assert L.avs_tune_set(7, 64, 32) == -1
        assert L.avs_tune_set(AVS_F64, 0, 32) == -2
        assert L.avs_tune_set(AVS_F64, 64, 48) == -3
        b = ctypes.c_int32()
        assert L.avs_tune_lookup(AVS_F64, 0, ctypes.byref(b)) == -2
        assert L.avs_tune_lookup(AVS_F64, 64, None) == -3
        # the serving-form entries of a table are what the library gets
        t = at.TuneTable()
        t.record("gfx950", np.float32, "svdvals", 2048, {"b": 64, "ms": 1.0})
        t.record("gfx950", np.float32, "band", 2048, {"b": 128, "ms": 1.0})
        assert at.push_to_library(t, arch="gfx950") == 1
        assert _lookup(AVS_F32, 100) == 64
        assert _lookup(AVS_F64, 4096) == 64  # cleared: defaults again
    finally:
        L.avs_tune_clear()


def test_packaged_table_is_well_formed():
    """Every shipped table parses and names built tile sizes only."""
    import glob
    import os
    for p in glob.glob(os.path.join(at.PKG_DIR, "tune_*.json")):
        t = at.TuneTable.load(p)
        for k, e in t.entries.items():
            arch, dt, form, bk = k.split("/")
            assert form in at.FORMS and np.dtype(dt) in (np.float32, np.float64)
            assert int(bk) == at.bucket(int(bk)) and e["b"] in at.TILES
