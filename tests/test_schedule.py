"""Host replay of the persistent executor's task DAG (CPU).

The device executor hands out tasks in the dispatch order the library builds
and runs each one once its per-tile completion counters reach the required
counts.  Here the same task list (avs_schedule_dump) is replayed with the
oracle's tile kernels (reference arithmetic, QRkernels.jl:3-197) under
randomized completion orders of a small worker pool:
  - every task becomes runnable (no deadlock; dispatch order is topological),
  - the flat tree reproduces the oracle's Blockbidiag! band bit for bit,
  - the hierarchical tree keeps sigma(A) and the band structure,
  - tiled_QR! keeps R = LAPACK's R up to row signs.
"""
import ctypes
import random

import numpy as np
import pytest
import scipy.linalg as sl

import avsvd
import oracle as O

_dp = ctypes.POINTER(ctypes.c_double)


class TileState:
    def __init__(self, a, b):
        n = a.shape[0]
        self.b, self.nt = b, n // b
        nt = self.nt
        self.t = np.zeros((nt * nt, b * b))
        for j in range(nt):
            for i in range(nt):
                self.t[j * nt + i] = a[i * b:(i + 1) * b, j * b:(j + 1) * b].reshape(-1, order="F")
        self.tau = {}

    def p(self, idx):
        return self.t[idx].ctypes.data_as(_dp)

    def dense(self, mode_band):
        b, nt = self.b, self.nt
        out = np.zeros((nt * b, nt * b))
        for j in range(nt):
            for i in range(nt):
                tile = self.t[j * nt + i].reshape((b, b), order="F")
                if mode_band:
                    tile = np.triu(tile) if i == j else (np.tril(tile) if j == i + 1 else 0 * tile)
                out[i * b:(i + 1) * b, j * b:(j + 1) * b] = tile
        return out


def _upper_view(tile_flat, b, trans):
    """logical upper triangle (of the tile or its adjoint) as a fresh tile."""
    m = tile_flat.reshape((b, b), order="F")
    lg = m.T if trans else m
    u = np.triu(lg)
    return np.asfortranarray(u.T if trans else u).reshape(-1, order="F").copy()


def run_task(S, t):
    L = O.lib()
    b = S.b
    tr = t["trans"]
    if t["kind"] == 0:  # panel
        tau = np.zeros(b)
        if t["mode"] == 0:
            L.or_geqrt(b, S.p(t["vtile"]), b, tr, tau.ctypes.data_as(_dp))
        elif t["mode"] == 1:
            L.or_tsqrt(b, S.p(t["vtile"]), b, S.p(t["btile"]), b, tr, tau.ctypes.data_as(_dp))
        else:
            bot = S.t[t["btile"]]
            tmp = _upper_view(bot, b, tr)
            L.or_tsqrt(b, S.p(t["vtile"]), b, tmp.ctypes.data_as(_dp), b, tr, tau.ctypes.data_as(_dp))
            # write back the logical upper triangle only (the lower part keeps
            # the domain head's GEQRT reflectors)
            m = bot.reshape((b, b), order="F")
            tm = tmp.reshape((b, b), order="F")
            mask = np.triu(np.ones((b, b), bool))
            if tr:
                mask = mask.T
            m[mask] = tm[mask]
            S.t[t["btile"]] = m.reshape(-1, order="F")
        S.tau[t["fac"]] = tau
        return
    tau = S.tau[t["fac"]]
    for q in range(t["ntiles"]):
        c2 = t["btile"] + q * t["cstep"]
        if t["mode"] == 0:
            L.or_unmqr(b, S.p(t["vtile"]), b, tr, tau.ctypes.data_as(_dp), S.p(c2), b, tr, b, 1)
        else:
            c1 = t["c1tile"] + q * t["cstep"]
            if t["mode"] == 1:
                v = S.t[t["vtile"]]
            else:
                v = _upper_view(S.t[t["vtile"]], b, tr)
            L.or_tsmqr(b, S.p(c1), b, S.p(c2), b, tr, v.ctypes.data_as(_dp), b, tr,
                       tau.ctypes.data_as(_dp), b, 1)


def replay(a, b, alg, tree, domain, workers, seed):
    S = TileState(a, b)
    tasks = avsvd.schedule_dump(S.nt, b, alg=alg, tree=tree, domain=domain, workers=8)
    cnt = np.zeros(S.nt * S.nt, dtype=np.int64)
    rng = random.Random(seed)
    nxt, running = 0, []
    while nxt < len(tasks) or running:
        while len(running) < workers and nxt < len(tasks):
            running.append(nxt)
            nxt += 1
        ready = [k for k in running if all(cnt[x] >= v for x, v in tasks[k]["deps"])]
        assert ready, "deadlock: no grabbed task can run"
        k = rng.choice(ready)
        running.remove(k)
        run_task(S, tasks[k])
        for x in tasks[k]["sigs"]:
            cnt[x] += 1
    return S, tasks


@pytest.mark.parametrize("b,nt", [(4, 5), (8, 4)])
def test_flat_tree_replay_matches_oracle_bitwise(b, nt):
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 31 * n)
    S, _ = replay(a, b, alg=0, tree=0, domain=0, workers=3, seed=1)
    ref = O.blockbidiag(a, b)
    np.testing.assert_array_equal(S.dense(True), ref)


@pytest.mark.parametrize("b,nt,domain,workers", [(4, 9, 2, 4), (4, 12, 3, 6), (8, 7, 2, 2), (4, 16, 4, 8)])
def test_hierarchical_tree_replay_keeps_sigma(b, nt, domain, workers):
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 17 * n + domain)
    S, tasks = replay(a, b, alg=0, tree=1, domain=domain, workers=workers, seed=domain)
    band = S.dense(True)
    assert O.sigma_err(sl.svdvals(band), sl.svdvals(a)) < 10 * n * np.finfo(float).eps
    assert any(t["mode"] == 2 for t in tasks)  # tree merges present


@pytest.mark.parametrize("tree,domain", [(0, 0), (1, 2)])
def test_tiled_qr_replay(tree, domain):
    b, nt = 4, 6
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 5 * n)
    S, _ = replay(a, b, alg=1, tree=tree, domain=domain, workers=4, seed=7)
    r = np.triu(S.dense(False))
    ref = sl.qr(a, mode="r")[0]
    np.testing.assert_allclose(np.abs(r), np.abs(ref), atol=1e-11)


def test_dependencies_are_topological():
    # every dependency refers to a write made by an earlier task in dispatch order
    for tree in (0, 1):
        tasks = avsvd.schedule_dump(12, 32, alg=0, tree=tree, domain=3, workers=16)
        writes = {}
        for k, t in enumerate(tasks):
            for x, v in t["deps"]:
                assert writes.get(x, 0) >= v, (k, x, v)
            for x in t["sigs"]:
                writes[x] = writes.get(x, 0) + 1


def test_schedule_stats_hier_shorter_than_flat():
    nf, ef = avsvd.schedule_stats(64, 128, tree=0, workers=256)
    nh, eh = avsvd.schedule_stats(64, 128, tree=1, domain=8, workers=256)
    assert nh > nf and eh < ef


def replay_phases(a, b, alg, tree, domain, seed):
    """Run the stream-ordered executor's phases with the oracle's tile
    kernels: phases in order, chains of a phase in random order, the ops of a
    chain in sequence, each update op over all its column tiles."""
    S = TileState(a, b)
    rng = random.Random(seed)
    phases = avsvd.phase_dump(S.nt, b, alg=alg, tree=tree, domain=domain)
    for kind, ncols, chains in phases:
        order = list(range(len(chains)))
        rng.shuffle(order)
        for c in order:
            for op in chains[c]:
                t = dict(op)
                if kind == 1:
                    t["ntiles"] = t["ncols"]
                run_task(S, t)
    return S, phases


@pytest.mark.parametrize("b,nt", [(4, 5), (8, 4)])
def test_flat_phases_match_oracle_bitwise(b, nt):
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 31 * n)
    S, _ = replay_phases(a, b, alg=0, tree=0, domain=0, seed=3)
    np.testing.assert_array_equal(S.dense(True), O.blockbidiag(a, b))


@pytest.mark.parametrize("b,nt,domain", [(4, 9, 2), (4, 12, 3), (8, 7, 2), (4, 16, 4)])
def test_hierarchical_phases_keep_sigma(b, nt, domain):
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 17 * n + domain)
    S, phases = replay_phases(a, b, alg=0, tree=1, domain=domain, seed=domain)
    assert O.sigma_err(sl.svdvals(S.dense(True)), sl.svdvals(a)) < 10 * n * np.finfo(float).eps
    # chains of one phase touch disjoint tiles (they run concurrently on the device)
    for kind, ncols, chains in phases:
        seen = set()
        for ch in chains:
            mine = set()
            for op in ch:
                if kind == 0:
                    mine |= {op["btile"]} if op["mode"] != 0 else set()
                else:
                    for q in range(op["ncols"]):
                        mine.add(op["btile"] + q * op["cstep"])
                        if op["mode"] != 0:
                            mine.add(op["c1tile"] + q * op["cstep"])
            assert not (mine & seen)
            seen |= mine


def test_tiled_qr_phases():
    b, nt = 4, 6
    n = b * nt
    a = O.gaussian_matrix(n, 20261203 + 5 * n)
    S, _ = replay_phases(a, b, alg=1, tree=1, domain=2, seed=7)
    np.testing.assert_allclose(np.abs(np.triu(S.dense(False))), np.abs(sl.qr(a, mode="r")[0]), atol=1e-11)
