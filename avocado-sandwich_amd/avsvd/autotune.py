"""Tile-size autotuner for the band reduction.

The reference fixes its tile size by hand and recompiles its kernels for it
(KAbasedSVD/src/QRkernelscompile.jl:1-4, tiledalgos.jl:262-267; the benchmark
scripts sweep it, runbenchmarks.jl:238-253).  Here every tile size in
{32, 64, 128} is compiled ahead of time, and this module times the candidate
configurations (tile size, merge-tree domain) on the device and keeps the
fastest per (GPU architecture, dtype, form, size bucket) in a JSON table:

  form "band"     device time of one band reduction (HIP events), the plain
                  single-matrix path;
  form "svdvals"  wall time of TiledMatrix.svdvals() (band reduction, band
                  D2H, host bulge chasing and dbdsdc), used by the serving loop;
  form "batch"    wall time of Batch.svdvals() for a set of mixed sizes,
                  keyed by the largest size.

The size bucket of n is the smallest power of two >= n; a lookup takes the
smallest bucket >= n, else the largest bucket of that (arch, dtype, form).
The table is read from $AVSVD_TUNE_FILE when set, else from the table shipped
with the package for the device's architecture (tune_<arch>.json).

    python -m avsvd.autotune plain --ns 4096 8192 16384 --dtypes float64 float32 --out tune.json
    python -m avsvd.autotune batch --count 4096 --out tune.json
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from ._lib import AVS_F32, AVS_F64, AvsError, check, lib

TILES = (32, 64, 128)
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
FORMS = ("band", "svdvals", "batch")


def bucket(n):
    """Smallest power of two >= n (n >= 1)."""
    n = int(n)
    if n < 1:
        raise ValueError("n must be positive")
    return 1 << (n - 1).bit_length()


def device_arch():
    """gfx architecture of device 0 (e.g. 'gfx950'), or None without a GPU.
    Asked of the library's HIP runtime, so that a process which has not
    imported torch does not load a second HIP runtime here."""
    buf = ctypes.create_string_buffer(64)
    if lib().avs_device_arch(0, buf, len(buf)) != 0:
        return None
    return buf.value.decode() or None


def _dtname(dtype):
    return np.dtype(dtype).name


def _key(arch, dtype, form, b):
    if form not in FORMS:
        raise ValueError("form must be one of %s" % (FORMS,))
    return "%s/%s/%s/%d" % (arch, _dtname(dtype), form, b)


class TuneTable:
    """Measured configurations keyed by (arch, dtype, form, size bucket)."""

    def __init__(self, entries=None, path=None):
        self.entries = dict(entries or {})
        self.path = path

    @classmethod
    def load(cls, path):
        with open(path) as f:
            doc = json.load(f)
        if doc.get("version") != 1:
            raise ValueError("%s: unknown tune table version" % path)
        return cls(doc["entries"], path)

    def save(self, path=None):
        path = path or self.path
        if not path:
            raise ValueError("no path to save the tune table to")
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({"version": 1, "entries": self.entries}, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
        self.path = path

    def record(self, arch, dtype, form, n, entry):
        self.entries[_key(arch, dtype, form, bucket(n))] = dict(entry)

    def buckets(self, arch, dtype, form):
        pre = "%s/%s/%s/" % (arch, _dtname(dtype), form)
        return sorted(int(k[len(pre):]) for k in self.entries if k.startswith(pre))

    def lookup(self, arch, dtype, form, n):
        """The entry for the smallest bucket >= n, else for the largest bucket;
        None when the table has nothing for (arch, dtype, form)."""
        bs = self.buckets(arch, dtype, form)
        if not bs:
            return None
        want = bucket(n)
        pick = next((b for b in bs if b >= want), bs[-1])
        return self.entries[_key(arch, dtype, form, pick)]


def default_table_path(arch=None):
    env = os.environ.get("AVSVD_TUNE_FILE")
    if env:
        return env
    arch = arch or device_arch()
    return os.path.join(PKG_DIR, "tune_%s.json" % arch) if arch else None


def load_table(path=None, arch=None):
    """The table at `path` (default: $AVSVD_TUNE_FILE, else the packaged table
    for the device's architecture); an empty table if there is none."""
    path = path or default_table_path(arch)
    if path and os.path.exists(path):
        return TuneTable.load(path)
    return TuneTable(path=path)


def default_tile(n, dtype, form):
    """Tile size without a measurement: the batched form 32; otherwise small
    tiles for small n (the host stage grows with b) and the largest tiles
    for n >= 16384 in FP64 band reductions."""
    if form == "batch":
        return 32
    if form == "svdvals":
        return 32 if n <= 2048 else 64
    if _dtname(dtype) == "float64" and n >= 16384:
        return 128
    return 32 if n <= 2048 else 64


def choose_tile(n, dtype, form="band", table=None, arch=None):
    """(b, domain) for an order-n matrix.  A tuned b that does not divide n is
    replaced by the fastest measured b that does (the plain path needs
    n % b == 0; the batched form pads, so any b fits).  domain 0 means the
    library default."""
    arch = arch or device_arch()
    table = table if table is not None else load_table(arch=arch)
    entry = table.lookup(arch, dtype, form, n) if arch else None
    divides = (lambda b: True) if form == "batch" else (lambda b: n % b == 0)
    if entry is not None:
        ranked = [entry["b"]] + [t[0] for t in sorted(entry.get("trials", []), key=lambda t: t[2])]
        for b in ranked:
            if divides(b):
                dom = entry.get("domain", 0) if b == entry["b"] else 0
                return int(b), int(dom)
    b = default_tile(n, dtype, form)
    if not divides(b):
        fits = [t for t in TILES if n % t == 0 and t <= b] or [t for t in TILES if n % t == 0]
        if not fits:
            raise AvsError(-4, "n=%d is not a multiple of any built tile size %s" % (n, TILES))
        b = max(fits)
    return b, 0


def push_to_library(table, arch=None):
    """Load the table's "svdvals" entries into the library's tile-size table,
    which the serving loop consults for requests without a fixed tile size."""
    arch = arch or device_arch()
    check(lib().avs_tune_clear())
    if not arch:
        return 0
    cnt = 0
    for dt, code in ((np.float32, AVS_F32), (np.float64, AVS_F64)):
        for bk in table.buckets(arch, dt, "svdvals"):
            e = table.lookup(arch, dt, "svdvals", bk)
            check(lib().avs_tune_set(code, bk, int(e["b"])))
            cnt += 1
    return cnt


# ---------------------------------------------------------------- measuring
def _best(trials):
    return min(trials, key=lambda t: t[2])


def tune_plain(n, dtype, form="band", tiles=TILES, domains=(0,), reps=2, seed=20261203, log=None):
    """Time every (tile, domain) with n % tile == 0 on a seeded N(0,1) matrix
    and return {"b", "domain", "ms", "trials": [[b, domain, ms], ...]}."""
    from .tiled import TiledMatrix, gaussian_matrix
    if form not in ("band", "svdvals"):
        raise ValueError("tune_plain measures the band or svdvals form")
    A0 = gaussian_matrix(n, seed, dtype=np.float64).astype(dtype, order="F")
    trials = []
    for b in tiles:
        if n % b or n < b:
            continue
        for dom in domains:
            A = A0.copy(order="F")
            T = TiledMatrix(A, b, b, domain=dom)
            best = float("inf")
            for _ in range(reps):
                T.reload()
                if form == "band":
                    check(lib().avs_blockbidiag(T._h, 0))
                    best = min(best, T.last_time_ms())
                else:
                    T.synchronize()
                    t0 = time.perf_counter()
                    T.svdvals()
                    best = min(best, 1e3 * (time.perf_counter() - t0))
            del T
            trials.append([b, dom, round(best, 4)])
            if log:
                log({"tune": form, "n": n, "dtype": _dtname(dtype), "b": b, "domain": dom, "ms": round(best, 4)})
    if not trials:
        raise AvsError(-4, "no built tile size divides n=%d" % n)
    b, dom, ms = _best(trials)
    return {"b": b, "domain": dom, "ms": ms, "trials": trials}


def tune_batch(ns, dtype=np.float32, tiles=(32, 64), reps=2, seed=20261203, log=None):
    """Time Batch.svdvals() of the matrices of orders `ns` (entries
    N(0, 1/n), seeded) for each tile size; returns the same dict as tune_plain."""
    from .batch import Batch
    from .tiled import gaussian_matrix
    mats = [gaussian_matrix(m, seed + i, scale=1.0 / np.sqrt(m), dtype=np.float64).astype(dtype, order="F")
            for i, m in enumerate(ns)]
    trials = []
    for b in tiles:
        best = float("inf")
        for _ in range(reps):
            bt = Batch(ns, dtype=dtype, b=b)
            for i, m in enumerate(mats):
                bt.set(i, m)
            t0 = time.perf_counter()
            bt.svdvals()
            best = min(best, 1e3 * (time.perf_counter() - t0))
            del bt
        trials.append([b, 0, round(best, 4)])
        if log:
            log({"tune": "batch", "count": len(ns), "nmax": max(ns), "dtype": _dtname(dtype), "b": b,
                 "ms": round(best, 4)})
    b, dom, ms = _best(trials)
    return {"b": b, "domain": dom, "ms": ms, "trials": trials}


def _emit(d):
    print(json.dumps(d), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m avsvd.autotune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("plain", help="tune the single-matrix path")
    p.add_argument("--ns", type=int, nargs="+", default=[1024, 2048, 4096, 8192, 16384])
    p.add_argument("--dtypes", nargs="+", default=["float64", "float32"])
    p.add_argument("--forms", nargs="+", default=["band", "svdvals"], choices=["band", "svdvals"])
    p.add_argument("--domains", type=int, nargs="+", default=[0])
    p.add_argument("--reps", type=int, default=2)
    p.add_argument("--svdvals-nmax", type=int, default=8192, help="largest n tuned for the svdvals form")
    p.add_argument("--out", required=True)
    q = sub.add_parser("batch", help="tune the batched form on the log-uniform {64..1024} mix")
    q.add_argument("--count", type=int, default=4096)
    q.add_argument("--dtype", default="float32")
    q.add_argument("--reps", type=int, default=2)
    q.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    arch = device_arch()
    if arch is None:
        print("autotune needs a GPU", file=sys.stderr)
        return 2
    table = TuneTable.load(a.out) if os.path.exists(a.out) else TuneTable(path=a.out)
    if a.cmd == "plain":
        for dt in a.dtypes:
            for n in a.ns:
                for form in a.forms:
                    if form == "svdvals" and n > a.svdvals_nmax:
                        continue
                    e = tune_plain(n, np.dtype(dt), form=form, domains=tuple(a.domains), reps=a.reps, log=_emit)
                    table.record(arch, dt, form, n, e)
                    _emit({"tuned": form, "n": n, "dtype": dt, "b": e["b"], "domain": e["domain"], "ms": e["ms"]})
                    table.save(a.out)
    else:
        rng = np.random.default_rng(20261203)
        ns = [int(2 ** (6 + int(np.floor(5 * u)))) for u in rng.random(a.count)]
        e = tune_batch(ns, np.dtype(a.dtype), reps=a.reps, log=_emit)
        table.record(arch, a.dtype, "batch", max(ns), e)
        _emit({"tuned": "batch", "count": a.count, "dtype": a.dtype, "b": e["b"], "ms": e["ms"]})
        table.save(a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
