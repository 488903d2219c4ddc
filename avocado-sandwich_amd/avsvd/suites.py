"""Benchmark suites for the workloads the band reduction is quoted on.

    python -m avsvd.suites plain     [--ns 4096 8192 16384] [--bs 32 64 128] [--dtypes f64 f32] [--kinds gauss geom]
    python -m avsvd.suites stage2    [--ns 4096 8192 16384] [--bs 32 64 128]
    python -m avsvd.suites precision [--n 16384] [--conds 1e2 1e4 1e6]
    python -m avsvd.suites batched   [--count 4096] [--plain-sample 128]
    python -m avsvd.suites serving   [--requests 48] [--budget-gb 64]
    python -m avsvd.suites ooc       [--n 16384] [--b 128] [--kswitches 32 64]
    python -m avsvd.suites qr        [--ns 1024 4096 16384] [--bs 64 128] [--dtypes f32 f64]
    python -m avsvd.suites phases    [--n 16384] [--dtype f32] [--bs 64 128] [--modes native split]

Every case prints one JSON line as soon as it finishes.  Inputs are synthetic
and seeded (Philox4x32-10 + Box-Muller from the library, seed 20261203 + case
index; prescribed spectra A = U diag(s) V^T with Haar U, V formed in FP64 on
the device).  The sigma reference is stated per line: the cached FP64 LAPACK
dgesdd vector of the same Gaussian matrix, the exact prescribed spectrum s, or
FP64 LAPACK computed here per matrix (batched / served forms).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import inputs
from ._lib import AVS_F32, AVS_F64, SCHED_STREAMS, TREE_HIER, UPD_NATIVE, UPD_SPLIT_BF16, Opts, check, dptr, lib
from .tiled import gaussian_matrix

SEED = 20261203
PEAK = {"f64": 78.6, "f32": 157.3}
GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "tests",
                      "golden")


def emit(d):
    print(json.dumps(d), flush=True)


def progress(msg):
    print(msg, file=sys.stderr, flush=True)


def sigma_err(sig, ref):
    s = np.sort(np.asarray(sig, dtype=np.float64))[::-1]
    r = np.sort(np.asarray(ref, dtype=np.float64))[::-1]
    return float(np.max(np.abs(s - r)) / r[0])


class Reducer:
    """One handle: device-resident input -> band reduction (timed) -> svdvals."""

    def __init__(self, n, b, dt, update_mode=UPD_NATIVE):
        o = Opts()
        lib().avs_default_opts(ctypes.byref(o))
        o.sched, o.tree, o.update_mode = SCHED_STREAMS, TREE_HIER, update_mode
        self.h = ctypes.c_void_p()
        self.n = n
        check(lib().avs_create(ctypes.byref(self.h), AVS_F64 if dt == "f64" else AVS_F32, n, b, ctypes.byref(o)))

    def close(self):
        if self.h.value:
            lib().avs_destroy(self.h)
            self.h = ctypes.c_void_p()

    def reduce(self, dptr_col_major, reps=2):
        ms = ctypes.c_double()
        best = None
        for _ in range(reps):
            check(lib().avs_set_matrix_device(self.h, ctypes.c_void_p(dptr_col_major), self.n))
            check(lib().avs_blockbidiag(self.h, 0))
            check(lib().avs_last_time_ms(self.h, ctypes.byref(ms)))
            best = ms.value if best is None else min(best, ms.value)
        return best

    def svdvals(self):
        out = np.empty(self.n)
        t0 = time.perf_counter()
        check(lib().avs_svdvals(self.h, dptr(out)))
        return out, time.perf_counter() - t0


def device_matrix(n, dt, kind, seed, cond):
    """(device tensor holding A column-major, reference sigma or None, reference label)."""
    import torch
    tdt = torch.float64 if dt == "f64" else torch.float32
    if kind == "gauss":
        a = gaussian_matrix(n, seed, dtype=np.float64)
        t = torch.from_numpy(a.T.astype(np.float64)).to("cuda:0").to(tdt).contiguous()
        path = os.path.join(GOLDEN, "sigma_gauss_n%d_s%d.npy" % (n, seed))
        ref = np.load(path) if os.path.exists(path) else None
        return t, ref, "FP64 LAPACK dgesdd of the FP64 matrix (cached)" if ref is not None else None
    t, s = inputs.geometric_matrix_device(n, seed, cond, dtype=np.float64 if dt == "f64" else np.float32)
    return t, s, "exact prescribed spectrum s_i = cond^(-(i-1)/(n-1)), cond=%g" % cond


def suite_plain(args):
    import torch
    for n in args.ns:
        for kind in args.kinds:
            for dt in args.dtypes:
                t, ref, rlabel = device_matrix(n, dt, kind, SEED, 1e6)
                torch.cuda.synchronize()
                for b in args.bs:
                    R = Reducer(n, b, dt)
                    ms = R.reduce(t.data_ptr())
                    sig, sv_s = R.svdvals()
                    R.close()
                    fl = 8.0 / 3.0 * n ** 3
                    d = {"suite": "plain", "n": n, "dtype": dt, "b": b, "input": kind,
                         "reduction_s": round(ms / 1e3, 5), "tflops": round(fl / ms / 1e9, 3),
                         "pct_of_peak": round(100 * fl / ms / 1e9 / PEAK[dt], 2),
                         "svdvals_s_host_stage_incl": round(sv_s, 3)}
                    if ref is not None:
                        d["sigma_err_over_sigma_max"] = sigma_err(sig, ref)
                        d["bar"] = 10 * n * float(np.finfo(np.float64 if dt == "f64" else np.float32).eps)
                        d["reference"] = rlabel
                    emit(d)
                del t
                torch.cuda.empty_cache()


def suite_stage2(args):
    """Band -> bidiagonal -> sigma after one FP64 band reduction, with the
    chase on the device (one launch per time of the pipelined schedule) and on
    the host (sweeps pipelined over the host threads): seconds of avs_svdvals
    after the reduction (band store or band D2H, chase, device bisection)."""
    import torch
    for n in args.ns:
        t, ref, rlabel = device_matrix(n, "f64", "gauss", SEED, 1e6)
        torch.cuda.synchronize()
        for b in args.bs:
            R = Reducer(n, b, "f64")
            ms = R.reduce(t.data_ptr(), reps=1)
            d = {"suite": "stage2", "n": n, "b": b, "reduction_s": round(ms / 1e3, 4)}
            sig = {}
            old = os.environ.get("AVSVD_STAGE2")
            try:
                for pol in ("device", "host"):
                    os.environ["AVSVD_STAGE2"] = pol
                    sig[pol], sec = R.svdvals()
                    d[pol + "_s"] = round(sec, 4)
            finally:
                if old is None:
                    os.environ.pop("AVSVD_STAGE2", None)
                else:
                    os.environ["AVSVD_STAGE2"] = old
            R.close()
            d["device_vs_host_sigma_err"] = sigma_err(sig["device"], sig["host"])
            if ref is not None:
                d["device_sigma_err_over_sigma_max"] = sigma_err(sig["device"], ref)
                d["reference"] = rlabel
            print(json.dumps(d), flush=True)
        del t


def suite_precision(args):
    import torch
    n = args.n
    for cond in args.conds:
        t64, s, _ = device_matrix(n, "f64", "geom", SEED + 1, cond)
        t32 = t64.to(torch.float32)
        torch.cuda.synchronize()
        for label, dt, mode, t in (("fp64", "f64", UPD_NATIVE, t64), ("fp32-native", "f32", UPD_NATIVE, t32),
                                   ("fp32-split-bf16", "f32", UPD_SPLIT_BF16, t32)):
            R = Reducer(n, args.b, dt, update_mode=mode)
            ms = R.reduce(t.data_ptr())
            sig, sv_s = R.svdvals()
            R.close()
            sig = np.sort(sig)[::-1]
            err = np.abs(sig - s)
            rel = err / s
            emit({"suite": "precision", "n": n, "b": args.b, "cond": cond, "form": label,
                  "reduction_s": round(ms / 1e3, 5), "tflops": round(8.0 / 3.0 * n ** 3 / ms / 1e9, 3),
                  "max_abs_err_over_sigma_max": float(err.max() / s[0]),
                  "median_rel_err": float(np.median(rel)), "max_rel_err_top_half": float(rel[: n // 2].max()),
                  "max_rel_err_bottom_decile": float(rel[-(n // 10):].max()),
                  "reference": "exact prescribed spectrum (FP64 formation of U diag(s) V^T)"})
        del t64, t32
        torch.cuda.empty_cache()


def _batched_inputs(count, seed):
    rng = np.random.default_rng(seed)
    ns = [int(2 ** (6 + math.floor(5 * u))) for u in rng.random(count)]
    ns = [min(x, 1024) for x in ns]
    mats = [None] * count

    def gen(i):
        n = ns[i]
        mats[i] = np.asfortranarray((gaussian_matrix(n, seed + 1 + i, dtype=np.float64) / math.sqrt(n))
                                    .astype(np.float32))
    with ThreadPoolExecutor(16) as ex:
        list(ex.map(gen, range(count)))
    return ns, mats


def _lapack_sigmas(mats):
    import scipy.linalg as sl
    from threadpoolctl import threadpool_limits
    out = [None] * len(mats)

    def job(i):
        out[i] = sl.svdvals(mats[i].astype(np.float64), check_finite=False)
        if i % 512 == 0:
            progress("lapack reference %d/%d" % (i, len(mats)))
    with threadpool_limits(1), ThreadPoolExecutor(16) as ex:
        list(ex.map(job, range(len(mats))))
    return out


def batched_run(count=4096, b=0, plain_sample=128, lapack_sample=0, seed=SEED):
    """The batched form: `count` FP32 matrices, n log-uniform in {64 .. 1024},
    entries N(0, 1/n), in one grouped launch per sweep; then the plain
    per-matrix FP64 GPU path, one matrix at a time, on a sample of the same
    mix.  sigma errors against FP64 LAPACK per matrix (every matrix, or a
    random sample of `lapack_sample`).  Returns one dict."""
    from .autotune import choose_tile
    from .batch import Batch
    from .tiled import TiledMatrix
    t0 = time.perf_counter()
    ns, mats = _batched_inputs(count, seed)
    gen_s = time.perf_counter() - t0
    flops = sum(8.0 / 3.0 * n ** 3 for n in ns)
    progress("batched: %d inputs in %.1f s" % (len(ns), gen_s))
    bt = Batch(ns, dtype=np.float32, b=b or None)
    bsel = bt.b
    for i, m in enumerate(mats):
        bt.set(i, m)
    t0 = time.perf_counter()
    sig = bt.svdvals()
    total_s = time.perf_counter() - t0
    red_ms = bt.last_time_ms()
    del bt
    rng = np.random.default_rng(seed + 7)
    chk = (list(range(len(ns))) if not lapack_sample or lapack_sample >= len(ns)
           else sorted(rng.choice(len(ns), size=lapack_sample, replace=False).tolist()))
    t0 = time.perf_counter()
    ref_list = _lapack_sigmas([mats[i] for i in chk])
    lap_s = time.perf_counter() - t0
    ref = dict(zip(chk, ref_list))
    errs = [sigma_err(sig[i], ref[i]) for i in chk]
    pick = rng.choice(len(ns), size=min(plain_sample, len(ns)), replace=False)
    t0 = time.perf_counter()
    perrs = []
    for c, i in enumerate(pick):
        if c % 16 == 0:
            progress("plain path %d/%d" % (c, len(pick)))
        A = np.asfortranarray(mats[i].astype(np.float64))
        bb = choose_tile(ns[i], np.float64, "svdvals")[0]
        T = TiledMatrix(A, bb, bb)
        s2 = T.svdvals()
        if i in ref:
            perrs.append(sigma_err(s2, ref[i]))
        del T
    plain_s = time.perf_counter() - t0
    # the plain path's rate over the batch's own size mix: mean time per
    # matrix of each size class in the sample, weighted by the batch's counts
    plain_rate = len(pick) / plain_s
    return {"count": len(ns), "b": bsel, "sizes": {str(k): ns.count(k) for k in sorted(set(ns))},
            "algorithmic_flops": flops, "band_reduction_s": round(red_ms / 1e3, 4),
            "band_tflops": round(flops / red_ms / 1e9, 3), "svdvals_total_s": round(total_s, 3),
            "svdvals_per_s": round(len(ns) / total_s, 2), "max_sigma_err_over_sigma_max": max(errs), "bar": 1e-5,
            "checked": len(chk),
            "reference": "FP64 LAPACK svdvals per matrix (%d matrices, %.1f s on 16 host threads)" % (len(chk), lap_s),
            "plain_fp64_one_at_a_time": {"sample": len(pick), "svdvals_per_s": round(plain_rate, 2),
                                         "max_sigma_err_over_sigma_max": max(perrs) if perrs else None},
            "speedup_over_plain_fp64": round((len(ns) / total_s) / plain_rate, 2)}


def suite_batched(args):
    d = batched_run(args.count, args.b, args.plain_sample)
    d["suite"] = "batched"
    emit(d)


def _serving_inputs(requests, nmax, seed):
    import torch
    rng = np.random.default_rng(seed + 11)
    reqs = []
    for i in range(requests):
        n = int(round(math.exp(rng.uniform(math.log(256), math.log(nmax)))))
        dt = np.float32 if rng.random() < 0.7 else np.float64
        kind = "gauss" if rng.random() < 0.5 else "geom"
        reqs.append((n, dt, kind, seed + 100 + i))
    mats, refs = [], []
    for c, (n, dt, kind, sd) in enumerate(reqs):
        progress("input %d/%d n=%d %s %s" % (c, len(reqs), n, np.dtype(dt).name, kind))
        if kind == "gauss":
            a = gaussian_matrix(n, sd, dtype=np.float64)
            ref = None
        else:
            a, ref = inputs.geometric_matrix_host(n, sd, 1e4, dtype=np.float64)
        mats.append(np.asfortranarray(a.astype(dt)))
        refs.append(ref)
    torch.cuda.empty_cache()
    return rng, reqs, mats, refs


def _serve(srv, mats, gaps):
    """Submits mats to the server at the given inter-arrival gaps (None: all
    at once) and waits for them; (outputs, wall seconds, p50 ms, p99 ms), the
    latencies being the server's per-request arrival-to-completion times."""
    ids = []
    t0 = time.perf_counter()
    tnext = t0
    for k, m in enumerate(mats):
        if gaps is not None:
            tnext += gaps[k]
            while time.perf_counter() < tnext:
                time.sleep(min(0.002, max(0.0, tnext - time.perf_counter())))
        ids.append(srv.submit(m))
    outs = []
    for c, i in enumerate(ids):
        outs.append(srv.wait(i))
        progress("served %d/%d" % (c + 1, len(ids)))
    wall = time.perf_counter() - t0
    lat = [o[1] for o in outs]
    return outs, wall, float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def serving_run(requests=48, nmax=16384, budget_gb=64, inflight=4, device_sigma=False, seed=SEED):
    """The served form: requests with n log-uniform in [256, nmax], 70 %% FP32,
    50/50 Gaussian / geometric (cond 1e4).  First the one-at-a-time plain GPU
    path on the same requests (and the same requests again in FP64, the plain
    per-matrix FP64 path), then the serving loop under Poisson arrivals at
    80 %% of the one-at-a-time capacity and at saturation (every request
    submitted at once), both on one long-lived server after a first
    saturated pass that fills its handle pool (reported as the cold pool).
    Returns one dict."""
    from .autotune import load_table, push_to_library
    from .server import served_tile
    from .tiled import TiledMatrix
    push_to_library(load_table())
    t0 = time.perf_counter()
    rng, reqs, mats, refs = _serving_inputs(requests, nmax, seed)
    gen_s = time.perf_counter() - t0

    def one_at_a_time(ms):
        lat = []
        t1 = time.perf_counter()
        for c, m in enumerate(ms):
            progress("one-at-a-time %d/%d (n=%d %s)" % (c, len(ms), m.shape[0], m.dtype.name))
            s0 = time.perf_counter()
            n = m.shape[0]
            b = served_tile(m.dtype, n)
            N = -(-n // b) * b
            A = np.zeros((N, N), dtype=m.dtype, order="F")
            A[:n, :n] = m
            T = TiledMatrix(A, b, b)
            T.svdvals()
            del T
            lat.append((time.perf_counter() - s0) * 1e3)
        return lat, len(ms) / (time.perf_counter() - t1)

    lat_plain, plain_rate = one_at_a_time(mats)
    _, plain64_rate = one_at_a_time([np.asfortranarray(m.astype(np.float64)) for m in mats])
    lam = 0.8 * plain_rate
    gaps = rng.exponential(1.0 / lam, size=len(mats))
    # one long-lived server, as a service runs it: a first pass at saturation
    # creates the pooled handles (the cold figures), then Poisson arrivals and
    # saturation again on the warm pool
    from .server import Server
    srv = Server(hbm_budget_gb=budget_gb, max_inflight=inflight, host_threads=16, device_sigma=device_sigma)
    cold_outs, cold_s, cold_p50, cold_p99 = _serve(srv, mats, None)
    outs, served_s, p50, p99 = _serve(srv, mats, gaps)
    sat_outs, sat_s, sat_p50, sat_p99 = _serve(srv, mats, None)
    srv.close()
    errs = [sigma_err(o[0], ref) for o, ref in zip(outs, refs) if ref is not None]
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0])
               for a, b, c in zip(outs, sat_outs, cold_outs))
    arr = np.cumsum(gaps)
    fin, lat_fifo = 0.0, []
    for a_s, svc_ms in zip(arr, lat_plain):
        fin = max(a_s, fin) + svc_ms / 1e3
        lat_fifo.append((fin - a_s) * 1e3)
    return {"requests": len(reqs), "n": [r[0] for r in reqs], "gen_s": round(gen_s, 1),
            "fp32_share": round(sum(r[1] == np.float32 for r in reqs) / len(reqs), 3),
            "hbm_budget_gb": budget_gb, "max_inflight": inflight, "device_sigma": bool(device_sigma),
            "saturated_cold_pool": {"p50_ms": round(cold_p50, 1), "p99_ms": round(cold_p99, 1),
                                    "solves_per_s": round(len(mats) / cold_s, 3)},
            "poisson": {"arrival_rate_per_s": round(lam, 3), "p50_ms": round(p50, 1),
                        "p99_ms": round(p99, 1), "solves_per_s": round(len(mats) / served_s, 3)},
            "saturated": {"p50_ms": round(sat_p50, 1), "p99_ms": round(sat_p99, 1),
                          "solves_per_s": round(len(mats) / sat_s, 3)},
            "one_at_a_time": {"p50_ms": round(float(np.percentile(lat_plain, 50)), 1),
                              "p99_ms": round(float(np.percentile(lat_plain, 99)), 1),
                              "solves_per_s": round(plain_rate, 3)},
            "one_at_a_time_fifo_same_arrivals": {"p50_ms": round(float(np.percentile(lat_fifo, 50)), 1),
                                                 "p99_ms": round(float(np.percentile(lat_fifo, 99)), 1)},
            "plain_fp64_one_at_a_time_solves_per_s": round(plain64_rate, 3),
            "saturated_over_plain_fp64": round((len(mats) / sat_s) / plain64_rate, 2),
            "saturated_cold_pool_over_plain_fp64": round((len(mats) / cold_s) / plain64_rate, 2),
            "results_same_cold_poisson_saturated": bool(same),
            "max_sigma_err_over_sigma_max_geometric_requests": max(errs) if errs else None,
            "reference": "exact prescribed spectrum (cond 1e4) for the geometric requests"}


def suite_serving(args):
    d = serving_run(args.requests, args.nmax, args.budget_gb, args.inflight, args.device_sigma)
    d["suite"] = "serving"
    emit(d)


def _streamed_bytes(n, b, ksteps, es):
    """Host<->device bytes of the streamed steps: every sweep moves its pivot
    and each of its units (nt - k tiles) up and down once."""
    nt = n // b
    tiles = 0
    for k in range(ksteps):
        tiles += (nt - k) * (nt - k)          # QR sweep: tile rows k..nt-1
        if k + 1 < nt:
            tiles += (nt - k - 1) * (nt - k)  # LQ sweep: tile columns k+1..nt-1
    return 2 * tiles * b * b * es


def suite_ooc(args):
    """OOC_Bidiag! (tiledalgos.jl:232-254) on a host Gaussian matrix against
    the in-core run with a host matrix (upload, reduction, download), and the
    singular values of the streamed result against the cached FP64 LAPACK
    vector of the same matrix."""
    from .tiled import OOC_Bidiag_, TiledMatrix, Blockbidiag_, _dense_to_band, bdsdc_values, bidiagonalize_band_chase
    n, b = args.n, args.b
    A0 = np.asfortranarray(gaussian_matrix(n, SEED, dtype=np.float64))
    path = os.path.join(GOLDEN, "sigma_gauss_n%d_s%d.npy" % (n, SEED))
    ref = np.load(path) if os.path.exists(path) else None
    A = A0.copy(order="F")
    t0 = time.perf_counter()
    T = TiledMatrix(A, b)
    Blockbidiag_(T)
    del T
    t_in = time.perf_counter() - t0
    progress("ooc: in-core with host matrix %.2f s" % t_in)
    for ks in args.kswitches:
        A = A0.copy(order="F")
        t0 = time.perf_counter()
        OOC_Bidiag_(A, kswitch=ks, tilesize=b)
        t_ooc = time.perf_counter() - t0
        ksteps = max(0, n // b - ks)
        moved = _streamed_bytes(n, b, ksteps, 8)
        d = {"suite": "ooc", "n": n, "b": b, "dtype": "f64", "kswitch": ks, "streamed_steps": ksteps,
             "ooc_s": round(t_ooc, 3), "in_core_host_matrix_s": round(t_in, 3),
             "streamed_bytes": moved, "streamed_gb_per_s": round(moved / t_ooc / 1e9, 2),
             "tflops": round(8.0 / 3.0 * n ** 3 / t_ooc / 1e12, 3),
             "device_tiles_held_streaming": 3 * (n // b)}
        if ref is not None:
            band = _dense_to_band(A, b)
            dd, ee = bidiagonalize_band_chase(band, n, b)
            d["sigma_err_over_sigma_max"] = sigma_err(bdsdc_values(dd, ee), ref)
            d["reference"] = "FP64 LAPACK dgesdd of the same matrix (cached)"
        emit(d)


class RocSolver:
    """The vendor QR: rocSOLVER ?geqrf / ?ormqr through the copies torch ships
    (built on the HIP runtime torch and this library share)."""

    def __init__(self, stream):
        import torch
        d = os.path.join(os.path.dirname(torch.__file__), "lib")
        self.blas = ctypes.CDLL(os.path.join(d, "librocblas.so"))
        self.sol = ctypes.CDLL(os.path.join(d, "librocsolver.so"))
        self.h = ctypes.c_void_p()
        if self.blas.rocblas_create_handle(ctypes.byref(self.h)) != 0:
            raise RuntimeError("rocblas_create_handle failed")
        if self.blas.rocblas_set_stream(self.h, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("rocblas_set_stream failed")
        vp, i = ctypes.c_void_p, ctypes.c_int
        for name in ("rocsolver_dgeqrf", "rocsolver_sgeqrf"):
            getattr(self.sol, name).argtypes = [vp, i, i, vp, i, vp]
        for name in ("rocsolver_dormqr", "rocsolver_sormqr"):
            getattr(self.sol, name).argtypes = [vp, i, i, i, i, i, vp, i, vp, vp, i]

    def geqrf(self, dt, n, pA, ptau):
        f = self.sol.rocsolver_dgeqrf if dt == "f64" else self.sol.rocsolver_sgeqrf
        if f(self.h, n, n, pA, n, ptau) != 0:
            raise RuntimeError("rocsolver geqrf failed")

    def ormqr_qt(self, dt, n, m, pA, ptau, pC):
        # side left (141), transpose (112): C <- Q^T C
        f = self.sol.rocsolver_dormqr if dt == "f64" else self.sol.rocsolver_sormqr
        if f(self.h, 141, 112, n, m, n, pA, n, ptau, pC, n) != 0:
            raise RuntimeError("rocsolver ormqr failed")

    def close(self):
        self.blas.rocblas_destroy_handle(self.h)


def _timed(fn, stream, reps):
    """median milliseconds of fn() on `stream` over reps runs (after one warmup)."""
    import torch
    fn()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def suite_qr(args):
    """tiled_QR! and myormqr! against the vendor geqrf / ormqr, as
    CAQR/src/benchmark.jl:14-47 compares with cuSOLVER: |R| and |Q^T C|
    differences relative to the vendor's, median times, vendor time / ours."""
    import torch
    # one stream for torch's copies, rocSOLVER and the handle (a handle given
    # stream 0 would make a stream of its own, unordered with the others)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        _qr_cases(args, stream)


def _qr_cases(args, stream):
    import torch
    rs = RocSolver(stream.cuda_stream)
    for dt in args.dtypes:
        tdt = torch.float64 if dt == "f64" else torch.float32
        for n in args.ns:
            g = torch.Generator(device="cuda:0").manual_seed(SEED + n)
            A0 = torch.randn(n, n, device="cuda:0", dtype=tdt, generator=g)  # column-major n x n in memory
            C0 = torch.randn(n, n, device="cuda:0", dtype=tdt, generator=g)
            tau = torch.zeros(n, device="cuda:0", dtype=tdt)
            Av, Cv = A0.clone(), C0.clone()
            t_geqrf = _timed(lambda: (Av.copy_(A0), rs.geqrf(dt, n, Av.data_ptr(), tau.data_ptr())), stream, args.reps)
            Av.copy_(A0)
            rs.geqrf(dt, n, Av.data_ptr(), tau.data_ptr())
            t_ormqr = _timed(lambda: (Cv.copy_(C0), rs.ormqr_qt(dt, n, n, Av.data_ptr(), tau.data_ptr(), Cv.data_ptr())),
                             stream, args.reps)
            Cv.copy_(C0)
            rs.ormqr_qt(dt, n, n, Av.data_ptr(), tau.data_ptr(), Cv.data_ptr())
            torch.cuda.synchronize()
            Rv = torch.triu(Av.T).abs()   # the memory is column-major: rows of Av.T are matrix rows
            QtCv = Cv.abs()
            t_copy = _timed(lambda: Av.copy_(A0), stream, args.reps)
            for b in args.bs:
                o = Opts()
                lib().avs_default_opts(ctypes.byref(o))
                o.stream = stream.cuda_stream
                h = ctypes.c_void_p()
                check(lib().avs_create(ctypes.byref(h), AVS_F64 if dt == "f64" else AVS_F32, n, b, ctypes.byref(o)))
                A = A0.clone()
                C = C0.clone()

                def ours_qr():
                    check(lib().avs_set_matrix_device(h, ctypes.c_void_p(A0.data_ptr()), n))
                    check(lib().avs_tiled_qr(h, 0))
                    check(lib().avs_get_matrix_device(h, ctypes.c_void_p(A.data_ptr()), n))

                def ours_q():
                    C.copy_(C0)
                    check(lib().avs_apply_q(h, ctypes.c_void_p(C.data_ptr()), n, n, 1))

                t_qr = _timed(ours_qr, stream, args.reps)
                kern = ctypes.c_double()
                check(lib().avs_last_time_ms(h, ctypes.byref(kern)))
                t_q = _timed(ours_q, stream, args.reps)
                torch.cuda.synchronize()
                R = torch.triu(A.T).abs()
                err_r = float(torch.linalg.norm((R - Rv).double()) / torch.linalg.norm(Rv.double()))
                err_q = float(torch.linalg.norm((C.abs() - QtCv).double()) / torch.linalg.norm(QtCv.double()))
                lib().avs_destroy(h)
                fl = 4.0 / 3.0 * n ** 3
                emit({"suite": "qr", "dtype": dt, "n": n, "b": b,
                      "geqrf_ms": round(t_qr, 3), "geqrf_kernels_ms": round(kern.value, 3),
                      "vendor_geqrf_ms": round(t_geqrf - t_copy, 3),
                      "vendor_over_ours_geqrf": round((t_geqrf - t_copy) / t_qr, 3),
                      "geqrf_tflops": round(fl / t_qr / 1e9, 3),
                      "rel_err_abs_R_vs_vendor": err_r,
                      "ormqr_ms": round(t_q - t_copy, 3), "vendor_ormqr_ms": round(t_ormqr - t_copy, 3),
                      "vendor_over_ours_ormqr": round((t_ormqr - t_copy) / max(t_q - t_copy, 1e-9), 3),
                      "rel_err_abs_QtC_vs_vendor": err_q,
                      "vendor": "rocSOLVER ?geqrf / ?ormqr (torch's copy), same stream",
                      "note": "our geqrf time includes packing A into tiles and writing R and the reflectors "
                              "back to an n x n array; the vendor's and both ormqr times are medians minus the "
                              "n x n device copy that restores their input"})
            del A0, C0, Av, Cv
            torch.cuda.empty_cache()
    rs.close()


def suite_phases(args):
    """Step time with the summed panel and update kernel times (per-launch
    events) on the stream-ordered executor, per tile size and update form."""
    import torch
    n = args.n
    dt = np.float64 if args.dtype == "f64" else np.float32
    A = gaussian_matrix(n, SEED, dtype=dt)
    dA = torch.from_numpy(A.T).to("cuda:0")
    torch.cuda.synchronize()
    for b in args.bs:
        for mode in args.modes:
            o = Opts()
            lib().avs_default_opts(ctypes.byref(o))
            o.sched, o.tree = SCHED_STREAMS, TREE_HIER
            o.update_mode = UPD_SPLIT_BF16 if mode == "split" else UPD_NATIVE
            o.phase_timing, o.reserve_cus, o.masked_modes = 1, args.reserve_cus, args.masked_modes
            h = ctypes.c_void_p()
            check(lib().avs_create(ctypes.byref(h), AVS_F64 if args.dtype == "f64" else AVS_F32, n, b, ctypes.byref(o)))
            for s in range(args.runs):
                check(lib().avs_set_matrix_device(h, ctypes.c_void_p(dA.data_ptr()), n))
                check(lib().avs_blockbidiag(h, 0))
                ms = ctypes.c_double()
                check(lib().avs_last_time_ms(h, ctypes.byref(ms)))
                pm, um, uf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
                pn, un = ctypes.c_int64(), ctypes.c_int64()
                check(lib().avs_phase_timing(h, ctypes.byref(pm), ctypes.byref(pn), ctypes.byref(um), ctypes.byref(un),
                                             ctypes.byref(uf)))
                if s == 0:
                    continue  # the first run builds the schedule
                emit({"suite": "phases", "n": n, "dtype": args.dtype, "b": b, "updates": mode, "run": s,
                      "step_ms": round(ms.value, 2), "panel_kernels_ms": round(pm.value, 2),
                      "panel_launches": pn.value, "update_kernels_ms": round(um.value, 2),
                      "update_launches": un.value, "update_tflops": round(uf.value / (um.value * 1e-3) / 1e12, 2)})
            lib().avs_destroy(h)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="suite", required=True)
    p = sub.add_parser("plain")
    p.add_argument("--ns", type=int, nargs="+", default=[4096, 8192, 16384])
    p.add_argument("--bs", type=int, nargs="+", default=[32, 64, 128])
    p.add_argument("--dtypes", nargs="+", default=["f64", "f32"])
    p.add_argument("--kinds", nargs="+", default=["gauss", "geom"])
    p = sub.add_parser("stage2")
    p.add_argument("--ns", type=int, nargs="+", default=[4096, 8192, 16384])
    p.add_argument("--bs", type=int, nargs="+", default=[32, 64, 128])
    p = sub.add_parser("precision")
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--b", type=int, default=128)
    p.add_argument("--conds", type=float, nargs="+", default=[1e2, 1e4, 1e6])
    p = sub.add_parser("batched")
    p.add_argument("--count", type=int, default=4096)
    p.add_argument("--b", type=int, default=0, help="tile size (0: the autotuner's table)")
    p.add_argument("--plain-sample", type=int, default=128)
    p = sub.add_parser("serving")
    p.add_argument("--requests", type=int, default=48)
    p.add_argument("--nmax", type=int, default=16384)
    p.add_argument("--budget-gb", type=float, default=64)
    p.add_argument("--inflight", type=int, default=4)
    p.add_argument("--device-sigma", action="store_true",
                   help="bidiagonals' singular values by bisection on the device in the serving loop")
    p = sub.add_parser("phases")
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--dtype", choices=["f64", "f32"], default="f32")
    p.add_argument("--bs", type=int, nargs="+", default=[64, 128])
    p.add_argument("--modes", nargs="+", default=["native", "split"])
    p.add_argument("--runs", type=int, default=3)
    p.add_argument("--reserve-cus", type=int, default=16)
    p.add_argument("--masked-modes", type=int, default=2)
    p = sub.add_parser("qr")
    p.add_argument("--ns", type=int, nargs="+", default=[1024, 2048, 4096, 8192, 16384])
    p.add_argument("--bs", type=int, nargs="+", default=[64, 128])
    p.add_argument("--dtypes", nargs="+", default=["f32", "f64"])
    p.add_argument("--reps", type=int, default=5)
    p = sub.add_parser("ooc")
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--b", type=int, default=128)
    p.add_argument("--kswitches", type=int, nargs="+", default=[32, 64])
    args = ap.parse_args(argv)
    {"plain": suite_plain, "stage2": suite_stage2, "precision": suite_precision, "batched": suite_batched,
     "serving": suite_serving, "ooc": suite_ooc, "qr": suite_qr,
     "phases": suite_phases}[args.suite](args)


if __name__ == "__main__":
    main()
