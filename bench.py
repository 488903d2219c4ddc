#!/usr/bin/env python3
"""bench.py -- n=16384 FP64 band reduction (Blockbidiag!) on one MI355X, plus
the batched and served svdvals forms and the reference's CPU path.

One step = one two-sided tiled QR/LQ reduction of a dense n x n matrix to
block-bidiagonal form (the reference's Blockbidiag!, KAbasedSVD/src/
tiledalgos.jl:17-34): repack the pristine column-major input (resident in HBM)
into the tile-major working copy, then run the tile DAG.  Algorithmic work per
step = (8/3) n^3 flops (the gebrd count; T factors, tree merges and padding
are not counted).  After the timed steps: the timed handle's singular values
(default svdvals path) against the cached FP64 dgesdd vector; the batched form
(4096 FP32 matrices of mixed order 64..1024: svdvals/s, sigma error); the
served form (Poisson and saturated request streams of mixed n and precision:
p50/p99 latency, solves/s, against the one-at-a-time plain path); the CPU
baseline (the reference's CPU plain path restated in C, stage by stage).
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384] [--b 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "avocado-sandwich_amd"))

METRIC = "n=16384 band-reduction TFLOP/s (% peak); mixed-size svdvals/s, p99 latency"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
SEED = 20261203


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ns=(1024, 4096), b=64):
    """The reference's CPU plain path (CUDAbasedSVD/SVD_GPU.jl:67-214 band
    reduction with LAPACK QR per cached block, then the block bulge chasing of
    KAbasedSVD/src/bulgechasing.jl:1-42 and dbdsdc, :50-53), restated in C in
    oracle/ and timed stage by stage on this host, with host FP64 dgesdd beside
    it.  The block loop runs on `threads` host threads, one BLAS thread each.
    value = (8/3) n^3 / band-stage seconds at the largest n."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    stages = {}
    for n in ns:
        a = oracle.gaussian_matrix(n, SEED + (1 if n == 1024 else 0))
        t0 = time.perf_counter()
        band = oracle.band_bidiagonal_cpu(a, b, threads)
        t1 = time.perf_counter()
        bd = oracle.bidiagonalize(band, b)
        t2 = time.perf_counter()
        sig = oracle.bdsdc(np.diag(bd).copy(), np.diag(bd, 1).copy())
        t3 = time.perf_counter()
        ref = oracle.svdvals_ref(a)
        t4 = time.perf_counter()
        stages[str(n)] = {"band_s": round(t1 - t0, 3), "band_tflops": round((8.0 / 3.0) * n ** 3 / (t1 - t0) / 1e12, 5),
                          "bulge_chasing_s": round(t2 - t1, 3), "bdsdc_s": round(t3 - t2, 4),
                          "gesdd_s": round(t4 - t3, 3), "sigma_err_vs_gesdd": oracle.sigma_err(sig, ref)}
        log("cpu baseline n=%d: %s" % (n, stages[str(n)]))
    top = stages[str(max(ns))]
    return {"value": top["band_tflops"], "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": "reference CPU plain path (BandBidiagonal!, CUDAbasedSVD/SVD_GPU.jl:67-214; LAPACK dgeqrt/"
                      "dgemqrt per cached block, block loop on %d threads with one BLAS thread each) restated in C, "
                      "FP64, b=%d, Gaussian n=%s; value = band stage at n=%d; bulge chasing "
                      "(bulgechasing.jl:1-42), dbdsdc and host dgesdd timed beside it" %
                      (threads, b, "/".join(str(x) for x in ns), max(ns)),
            "stages": stages, "host_cpus": os.cpu_count()}


def panel_rate(n, b, tree, domain, dtype, panel_ms):
    """Achieved HBM rate of the panel kernels: algorithmic bytes of every panel
    operation of one reduction (GEQRT reads and writes its tile, 2 b^2 elements;
    TSQRT and the merges their two tiles' R triangle and bottom tile,
    2 (b(b+1)/2 + b^2) elements) over the panel launches' summed duration per
    reduction.  The panels are latency-bound; the rate is low by nature."""
    import avsvd
    elt = 8 if dtype == "f64" else 4
    ops = {0: 0, 1: 0, 2: 0}
    for kind, _, chains in avsvd.phase_dump(n // b, b, 0, 0, tree, domain):
        if kind == 0:
            for ch in chains:
                for op in ch:
                    ops[op["mode"]] += 1
    nbytes = (ops[0] * 2 * b * b + (ops[1] + ops[2]) * 2 * (b * (b + 1) // 2 + b * b)) * elt
    return {"bound": "latency", "algorithmic_bytes_per_step": nbytes, "panel_ms_per_step": round(panel_ms, 3),
            "achieved_gbps": round(nbytes / panel_ms / 1e6, 2) if panel_ms > 0 else None, "peak_gbps": 8000.0,
            "ops": {"ge": ops[0], "ts": ops[1], "tt": ops[2]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--b", type=int, default=128)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--sched", choices=["streams", "host"], default="streams")
    ap.add_argument("--tree", choices=["hier", "flat"], default="hier")
    ap.add_argument("--domain", type=int, default=0)
    ap.add_argument("--reserve-cus", type=int, default=0,
                    help="CUs (the same number per XCD) a CU-masked update stream leaves to the panels (0: none)")
    ap.add_argument("--masked-modes", type=int, default=2,
                    help="update modes on the CU-masked stream: bit 0 GEQRT's, 1 TSQRT's, 2 merges' (0 = all)")
    ap.add_argument("--cpu-ns", type=int, nargs="+", default=[1024, 4096])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check-n", type=int, default=16384,
                    help="sigma check against the cached FP64 dgesdd vector: the timed handle's own result when "
                         "equal to --n, else a separate run at this size (0 = skip)")
    ap.add_argument("--batched-count", type=int, default=4096, help="batched form: matrices (0 = skip)")
    ap.add_argument("--serving-requests", type=int, default=24, help="served form: requests (0 = skip)")
    ap.add_argument("--serving-nmax", type=int, default=16384)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (a known-traffic kernel for scaling "
                         "hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the multi-GPU sharding of the reduction is a design, "
              "not built); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    import avsvd
    from avsvd._lib import Opts, check, dptr, lib

    L = lib()
    n, b = args.n, args.b
    dt = np.float64 if args.dtype == "f64" else np.float32
    code = avsvd.AVS_F64 if dt == np.float64 else avsvd.AVS_F32
    dev = torch.device("cuda:0")

    log("generating n=%d %s Gaussian (seed %d)" % (n, args.dtype, SEED))
    A = avsvd.gaussian_matrix(n, SEED, dtype=dt)
    fro = float(np.linalg.norm(A.astype(np.float64)))
    dA = torch.from_numpy(A.T).to(dev)  # row-major (n,n) of A^T == column-major A
    torch.cuda.synchronize()

    o = Opts()
    L.avs_default_opts(ctypes.byref(o))
    o.sched = avsvd.SCHED_STREAMS if args.sched == "streams" else avsvd.SCHED_HOST
    o.tree = avsvd.TREE_HIER if args.tree == "hier" else avsvd.TREE_FLAT
    o.domain = args.domain
    o.wait_ms = 120000
    # per-launch timing events (for the roofline) only in the last timed step:
    # they add launch latency to the panel chain (n=16384 FP64 b=128: 374 vs
    # 356 ms per reduction)
    o.phase_timing = 0
    if args.sched == "streams":
        # optionally the updates of some modes on a stream that leaves
        # reserve_cus CUs to the panels (bit-identical bands).  With the
        # blocked panel every update on all CUs is fastest at n=16384 FP64
        # b=128: 391.6 ms against 399.5-400.1 ms with 8-24 CUs left to the
        # panels beside the TSQRT updates
        o.reserve_cus = args.reserve_cus
        o.masked_modes = args.masked_modes
    h = ctypes.c_void_p()
    check(L.avs_create(ctypes.byref(h), code, n, b, ctypes.byref(o)))
    pdA = ctypes.c_void_p(dA.data_ptr())

    def step():
        check(L.avs_set_matrix_device(h, pdA, n))
        check(L.avs_blockbidiag(h, 0))

    ms = ctypes.c_double()
    t_build = time.perf_counter()
    for w in range(args.warmup):
        step()
        check(L.avs_synchronize(h))
        check(L.avs_last_time_ms(h, ctypes.byref(ms)))
        log("warmup %d: kernel %.1f ms (first includes schedule build %.1f s)" %
            (w, ms.value, time.perf_counter() - t_build))
    kernel_ms = []
    upd = {"ms": 0.0, "launches": 0, "flops": 0.0, "panel_ms": 0.0, "panel_launches": 0, "steps": 1}
    check(L.avs_synchronize(h))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        if s == args.steps - 1 and args.sched == "streams":
            check(L.avs_set_phase_timing(h, 1))
        step()
        check(L.avs_last_time_ms(h, ctypes.byref(ms)))  # waits for this step's kernels
        kernel_ms.append(ms.value)
    check(L.avs_synchronize(h))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s, k in enumerate(kernel_ms):
        log("step %d: %.1f ms" % (s, k))
    if args.sched == "streams":
        # per-launch HIP events of the last timed step (the only step that
        # records them; read after the timed region, so the host-side readout
        # of thousands of events is not counted in the step time)
        pm, um, uf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        pn, un = ctypes.c_int64(), ctypes.c_int64()
        check(L.avs_phase_timing(h, ctypes.byref(pm), ctypes.byref(pn), ctypes.byref(um), ctypes.byref(un),
                                 ctypes.byref(uf)))
        upd = {"ms": um.value, "launches": un.value, "flops": uf.value, "panel_ms": pm.value,
               "panel_launches": pn.value, "steps": 1}
        log("last step: update kernels %.1f ms in %d launches, panel kernels %.1f ms in %d" %
            (um.value, un.value, pm.value, pn.value))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    flops = (8.0 / 3.0) * float(n) ** 3
    value = flops / (ms_per_step * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if dt == np.float64 else PEAK_FP32_TFLOPS
    if upd["launches"]:
        # dominant kernel: chain_update_kernel; algorithmic flops per launch / mean launch duration
        avg_ms = upd["ms"] / upd["launches"]
        flops_per_launch = upd["flops"] / upd["launches"]
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        kname = "chain_update_kernel<%s,%d>" % ("double" if dt == np.float64 else "float", b)
    else:
        avg_ms = sum(kernel_ms) / len(kernel_ms)
        flops_per_launch = flops
        achieved = flops / (avg_ms * 1e-3) / 1e12
        kname = "whole reduction (host-enqueued launches)"

    # size-independent check on the full-size result: ||band||_F == ||A||_F
    band = np.zeros((b + 1, n), order="F")
    check(L.avs_get_band(h, dptr(band), b + 1))
    fro_err = abs(float(np.linalg.norm(band)) - fro) / fro
    del band

    sigma_check = None
    ref_path = os.path.join(ROOT, "tests", "golden", "sigma_gauss_n%d_s%d.npy" % (args.check_n, SEED))
    if args.check_n and os.path.exists(ref_path):
        ref = np.load(ref_path)
        t_s = time.perf_counter()
        if args.check_n == n:
            # the timed handle's own band (same options as the timed steps), then the default svdvals path
            sig = np.empty(n)
            check(L.avs_svdvals(h, dptr(sig)))
            what = "the timed handle's band, then the default svdvals path"
        else:
            a_chk = avsvd.gaussian_matrix(args.check_n, SEED, dtype=dt)
            T = avsvd.TiledMatrix(a_chk, b, b, sched=o.sched, tree=o.tree, domain=args.domain,
                                  reserve_cus=o.reserve_cus, masked_modes=o.masked_modes)
            sig = T.svdvals()
            del T
            what = "a separate handle with the same options"
        t_s = time.perf_counter() - t_s
        err = float(np.max(np.abs(np.sort(sig)[::-1] - ref)) / ref[0])
        sigma_check = {"n": args.check_n, "max_abs_err_over_sigma_max": err, "bar": 1e-12 if args.dtype == "f64" else 1e-5,
                       "vs": "FP64 LAPACK dgesdd of the same seeded matrix (cached)", "what": what,
                       "svdvals_seconds": round(t_s, 3)}
        log("sigma check n=%d: %.3e (%.2f s)" % (args.check_n, err, t_s))
    L.avs_destroy(h)
    del dA
    torch.cuda.empty_cache()

    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.empty(nb // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB device-to-device copy done" % args.copy_calibration)
        del src, dst

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_blockbidiag_n%d_%s_b%d.json" % (n, args.dtype, b))
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    batched = serving = None
    from avsvd import suites
    if args.batched_count:
        log("batched form: %d FP32 matrices" % args.batched_count)
        try:
            batched = suites.batched_run(args.batched_count, 0, plain_sample=32, lapack_sample=256)
        except Exception as e:  # the headline line is still printed
            batched = {"error": repr(e)}
        torch.cuda.empty_cache()
    if args.serving_requests:
        log("served form: %d requests" % args.serving_requests)
        try:
            serving = suites.serving_run(args.serving_requests, args.serving_nmax, budget_gb=64, inflight=4,
                                         device_sigma=True)
        except Exception as e:
            serving = {"error": repr(e)}
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(tuple(args.cpu_ns), 64)
        except Exception as e:
            cpu = {"error": repr(e)}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: iid N(0,1) entries, Philox4x32-10 seed %d" % SEED,
        "config": {"workload": "Blockbidiag! n=%d %s tile b=%d, Gaussian input resident in HBM" %
                               (n, args.dtype.upper(), b),
                   "n": n, "b": b, "sched": args.sched, "tree": args.tree, "domain": args.domain or "auto",
                   "reserved_cus": args.reserve_cus if args.sched == "streams" else 0,
                   "masked_update_modes": args.masked_modes if args.sched == "streams" else 0,
                   "parallelism": "single GPU"},
        "pct_of_peak": round(100.0 * value / peak, 2),
        "seconds_per_reduction": round(ms_per_step / 1e3, 4),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": kname,
                     "avg_launch_ms": round(avg_ms, 5), "algorithmic_flops_per_launch": flops_per_launch,
                     "launches": upd["launches"], "kernel_share_of_step": round(upd["ms"] / kernel_ms[-1], 4)
                     if upd["launches"] else None,
                     "panel_ms_per_step": round(upd["panel_ms"], 3),
                     "step_kernel_ms": [round(x, 3) for x in kernel_ms],
                     "launch_events": "per-launch HIP events on each launch's own stream, last timed step only"},
        "panels": panel_rate(n, b, o.tree, args.domain, args.dtype, upd["panel_ms"]),
        "band_fro_rel_err": fro_err,
        "sigma_check": sigma_check,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
