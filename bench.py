"""Benchmark of the amalgamated-likelihood DP on one MI355X.

One step = one log-likelihood + gradient evaluation of the whole batch at a
parameter point: slice tables (k_tables), forward DP of every family
(k_forward), reverse sweep (k_backward) and the table adjoint / batch sum
(k_tail).  The workload is the large-batch configuration (config4: 50-taxon
species tree with 4 WGDs, Δt = 0.01, 10,000 families of 4–200 leaves), with
synthetic, seeded data; inputs are device-resident before timing starts.

Printed JSON line: families/s of batched FP32-scaled logL+grad, with its
error against the FP64-scaled form, the FP64 plain form launched per family
for comparison, the roofline of the dominant kernel and the CPU oracle's
throughput on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config4]
                    [--no-cpu] [--no-per-family] [--cpu-seconds S]
                    [--no-sampler] [--no-tune] [--tune-reps R] [--tune-cache FILE]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per evaluation of `kernel` (all its size-class launches) on this
    workload, from the FETCH_SIZE/WRITE_SIZE profile committed under profiles/
    (copy-calibrated), or None when the profile is absent."""
    path = os.path.join(ROOT, "profiles", "config4_fp32_gradient_hbm_traffic.md")
    sym = {"k_backward": "`k_backward<float,true,true,false>`",
           "k_forward": "`k_forward<float,true,true,true,false>`"}[kernel]
    try:
        for line in open(path, encoding="utf-8"):
            if line.startswith("| " + sym + " |"):
                return float(line.strip().strip("|").split("|")[-1]) * 1e6
    except OSError:
        pass
    return None


def algorithmic_bytes(st, s_bytes, x_fwd=0, x_bwd=0):
    """Algorithmic HBM bytes of one evaluation: each ℓ cell written once and
    read once in the forward (2·s per cell), 16 B per split (two int32 ids and
    an fp64 p), 4 B per compatible clade, the exponents X stored by the scaled
    forms and the 8-byte result; the gradient adds the ℓ re-read, the adjoint
    write and read (3·s per cell) and the split data again."""
    fwd = 2 * s_bytes * st["cells"] + 16 * st["splits"] + 4 * st["sum_compat"] + x_fwd + 8
    grad = fwd + 3 * s_bytes * st["cells"] + 16 * st["splits"] + x_bwd
    return fwd, grad


def copy_probe(nbytes=1 << 28):
    """One device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE in a
    rocprofv3 --pmc run give the counters' scale (KiB read per byte moved)."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes device-to-device")


def sampler(w, synth, workload="config5", n_points=256, n_families=None, tune_cache=None, tune=True):
    """Sampler-style loop (config5: 20 taxa + 5 WGDs, 5,000 families):
    device-resident CCDs, parameter points drawn as log-normal perturbations
    of the truth, each evaluated (logL + gradient, FP32-scaled, batched)
    back to back as HMC leapfrog steps would.  Latency is host call to host
    result, parameters in and gradient out included."""
    t0 = time.time()
    wl = synth.workload(workload, n_families=n_families)
    model = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    ccds = w.ccds_from_texts(wl.texts, model)
    batch = w.device_batch(model, ccds)
    n_fam = batch.n
    points = synth.perturbed(wl.rates, n_points)
    models = [model(r) for r in points]
    log(f"sampler workload {workload}: {n_fam} families ready in {time.time() - t0:.1f}s")
    tuned = None
    if tune:
        model._sync()
        tuned = batch.autotune("fp32", gradient=True, reps=5, cache=tune_cache)
        log(f"sampler autotune: {tuned}")
    for m in models[:3]:
        w.gradient(m, ccds, mode="fp32")
    lat = []
    t0 = time.perf_counter()
    for m in models:
        t = time.perf_counter()
        w.gradient(m, ccds, mode="fp32")
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    # the same loop with bit-identical gradients (reversible trajectories)
    batch.set_deterministic(True)
    for m in models[:3]:
        w.gradient(m, ccds, mode="fp32")
    lat_det = []
    for m in models[:64]:
        t = time.perf_counter()
        w.gradient(m, ccds, mode="fp32")
        lat_det.append(time.perf_counter() - t)
    batch.set_deterministic(False)
    lat_det = np.array(lat_det) * 1e3
    out = {"workload": f"{wl.name}: {n_fam} families, {len(points)} parameter points back to back, "
                       "logL+gradient FP32-scaled, device-resident CCDs",
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "evals_per_s": len(points) / total, "families_per_s": n_fam * len(points) / total,
           "deterministic_p50_ms": float(np.percentile(lat_det, 50)),
           "deterministic_p99_ms": float(np.percentile(lat_det, 99)),
           "tuning": tuned}
    log(f"sampler: p50 {out['p50_ms']:.3f} ms, p99 {out['p99_ms']:.3f} ms, {out['evals_per_s']:.1f} evals/s; "
        f"deterministic p50 {out['deterministic_p50_ms']:.3f} ms")
    return out


def cpu_threads():
    """Host threads the CPU baseline uses: every hardware thread of this
    process's CPU set, at most 16 (one GPU's share of a shared 8-GPU host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(wl, n_fam, threads, seconds, k=None):
    """The oracle (C++ FP64 restatement of the reference: FP64 logL plus a
    ForwardDiff-style gradient in chunks of 12 duals) on every k-th family,
    k chosen so that the gradient pass takes about `seconds` unless given;
    families/s."""
    from oracle.oracle import OracleModel, DLWGD as ODLWGD
    om = OracleModel(wl.newick, [], wl.dt)
    r = wl.rates
    orr = ODLWGD(list(r.λ), list(r.μ), list(r.q), r.η)
    fixed = k is not None
    k = k if fixed else max(1, n_fam // 64)
    while True:
        idx = list(range(0, n_fam, k))
        oc = om.ccds([wl.texts[i] for i in idx])
        t = time.perf_counter()
        om.gradient(oc, orr, nthreads=threads)
        dt = time.perf_counter() - t
        if fixed or dt * 1.5 > seconds or k == 1:
            break
        k = max(1, int(k * dt / seconds))
    return dict(value=len(idx) / dt, unit="families/s", cores=threads, kind="port", every=k,
                sample=f"logL+ForwardDiff-style gradient (chunks of 12 duals) of {len(idx)} of the {n_fam} "
                       f"families (every {k}th), oracle C++ FP64 restatement of the reference, {dt:.1f} s on "
                       f"{threads} threads")


def cpu_calibration():
    """The reference's own benchmark workload (test/runtests.jl:21-42: the 12
    example-1 families on the 9-taxon tree with 2 WGDs, Δt = 0.05,
    RootCondition) timed on the oracle at 2 threads, to be read beside the
    reference's 220.969 µs (logpdf) and 2.723 ms (ForwardDiff gradient) on
    its authors' 2-thread machine."""
    import glob
    from oracle.oracle import OracleModel, DLWGD as ODLWGD
    extree = ("((MPOL:4.752,PPAT:4.752):0.292,(SMOE:4.457,(((OSAT:1.555,(ATHA:0.5548,CPAP:0.5548):1.0002):0"
              ".738,ATRI:2.293):1.225,(GBIL:3.178,PABI:3.178):0.34):0.939):0.587);")
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "example-1", "*.ale")))
    om = OracleModel(extree, [("wgd_1", "ATHA", "ATHA"), ("wgd_2", "ATHA", "ATRI")], 0.05)
    oc = om.ccds([open(f).read() for f in files])
    r = ODLWGD([1.0] * 17, [1.0] * 17, [0.2, 0.1], 0.9)
    v = om.logpdf(oc, r, nthreads=2)

    def median_time(fn, reps):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))
    t_l = median_time(lambda: om.logpdf(oc, r, nthreads=2), 200)
    t_g = median_time(lambda: om.gradient(oc, r, nthreads=2), 50)
    return {"workload": "KAT-2: 12 example-1 families, 9-taxon tree + 2 WGDs, dt=0.05 (test/runtests.jl:21-42)",
            "threads": 2, "logL": v, "logpdf_us": 1e6 * t_l, "gradient_ms": 1e3 * t_g,
            "reference_logpdf_us": 220.969, "reference_gradient_ms": 2.723,
            "note": "reference figures from its test file, on its authors' unnamed 2-thread machine"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="config4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-family", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--families", type=int, default=None)
    ap.add_argument("--no-sampler", action="store_true", help="skip the config5 sampler-loop leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the hand-picked launch shape (no autotuning)")
    ap.add_argument("--tune-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "whale_amd", "tuning.txt"),
                    help="autotuner cache file (empty: search every time)")
    ap.add_argument("--sampler-points", type=int, default=256)
    ap.add_argument("--tune-reps", type=int, default=5, help="timed replays per candidate shape in the search")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one device-to-device copy of 256 MiB (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (families would be sharded across "
              "GPUs with one all-reduce of the gradient). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    w = load_package()
    from whale_jl_amd import _lib, synth
    from whale_jl_amd._lib import check, dblp, ptr
    L = _lib.lib()
    t0 = time.time()
    wl = synth.workload(args.workload, n_families=args.families)
    log(f"workload {args.workload}: {len(wl.texts)} families generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    model = w.WhaleModel(wl.rates, w.readnw(wl.newick), wl.dt)
    ccds = w.ccds_from_texts(wl.texts, model)
    batch = w.device_batch(model, ccds)
    st = batch.stats()
    log(f"parsed and planned in {time.time() - t0:.1f}s: {st}")
    n_fam = st["families"]
    ctx = w.context()
    h = model.s.handle()
    lam, mu, q, eta, p, _ = model.node_params()
    q = np.ascontiguousarray(q if len(q) else np.zeros(1))
    nn, nw = model.s.n, model.s.n_wgd
    fam = np.zeros(n_fam)
    dl, dm, dq = np.zeros(nn), np.zeros(nn), np.zeros(max(1, nw))
    deta, tot = C.c_double(), C.c_double()

    def step(mode, launch=_lib.BATCHED, grad=True):
        check(L.whale_model_set_params(h, ptr(lam, dblp), ptr(mu, dblp), ptr(q, dblp), len(q), eta, ptr(p, dblp)),
              "set_params")
        if grad:
            check(L.whale_logpdf_grad(batch.h, mode, launch, ptr(fam, dblp), C.byref(tot), ptr(dl, dblp),
                                      ptr(dm, dblp), ptr(dq, dblp), C.byref(deta)), "logpdf_grad")
        else:
            check(L.whale_logpdf(batch.h, mode, launch, ptr(fam, dblp), C.byref(tot)), "logpdf")
        return tot.value, fam.copy(), np.concatenate([dl, dm, dq[:nw], [deta.value]])

    # ---- accuracy: FP32-scaled vs FP64-scaled (no underflow) and FP64 plain
    v64s, f64s, g64s = step(_lib.FP64_SCALED)
    v64, f64, g64 = step(_lib.FP64_PLAIN)
    v32, f32, g32 = step(_lib.FP32_SCALED)
    n_fallback = batch.fallbacks()
    fin = np.isfinite(f64s)
    rel_fam = float(np.max(np.abs(f32[fin] - f64s[fin]) / np.abs(f64s[fin])))
    rel_tot = abs(v32 - v64s) / abs(v64s)
    gerr = float(np.max(np.abs(g32 - g64s)) / np.max(np.abs(g64s)))
    n_underflow = int(np.sum(~np.isfinite(f64)))
    log(f"logL fp64s {v64s:.6f} fp64 {v64:.6f} fp32 {v32:.6f}; fp32 rel err family max {rel_fam:.2e} "
        f"total {rel_tot:.2e} grad {gerr:.2e}; fp64-plain underflowed families {n_underflow}; "
        f"fp32 families recomputed in fp64 {n_fallback}")

    # ---- launch shape: the autotuner's search on this batch (or its cached result)
    tune = None
    if not args.no_tune:
        cache = args.tune_cache or None
        if cache:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
        tune = batch.autotune("fp32", gradient=True, reps=args.tune_reps, cache=cache)
        tune["tuning_of"] = "class LDS bounds (bytes), stream layout, waves per CU; ms are host-timed graph replays"
        log(f"autotune: {tune}")
    for r in batch.launches("fp32"):  # the shape the timed steps run
        if not r["families"]:
            continue
        log("launch fp32 {kernel} class {class}: {families} families, {per_workgroup}/workgroup, tables in LDS "
            "{tables_in_lds}, grid {grid}, LDS {lds_bytes} B ({shared_lds_bytes} shared, {family_lds_bytes} per "
            "family)".format(**r))

    # ---- timed: batched FP32-scaled logL + gradient (captured HIP graph replays)
    for _ in range(args.warmup):
        step(_lib.FP32_SCALED)
    t_wall = time.perf_counter()
    check(L.whale_ctx_record(ctx, 0), "record")  # HIP events on the library's stream
    for _ in range(args.steps):
        step(_lib.FP32_SCALED)
    check(L.whale_ctx_record(ctx, 1), "record")
    ms_c = C.c_double()
    check(L.whale_ctx_elapsed(ctx, 0, 1, C.byref(ms_c)), "elapsed")
    wall = time.perf_counter() - t_wall
    ms_step = ms_c.value / args.steps
    value = n_fam / (ms_step / 1e3)
    # per-kernel breakdown: the same steps launched eagerly with events between the kernels
    L.whale_ctx_set_timing(ctx, 1)
    ms5 = np.zeros(5)
    runs = []
    nk = max(3, args.steps // 2)
    for _ in range(nk):
        step(_lib.FP32_SCALED)
        L.whale_last_timings(ctx, ptr(ms5, dblp))
        runs.append(ms5.copy())
    L.whale_ctx_set_timing(ctx, 0)
    kern = np.median(np.array(runs), axis=0)  # [tables, forward, backward, tail, total]
    log("eager per-phase ms (tables fwd bwd tail), each evaluation: " +
        "; ".join(" ".join(f"{x:.3f}" for x in r[:4]) for r in runs))
    log(f"fp32-scaled batched: {ms_step:.3f} ms/step (wall {1e3 * wall / args.steps:.3f}), {value:.4g} families/s; "
        f"kernels ms: tables {kern[0]:.3f} fwd {kern[1]:.3f} bwd {kern[2]:.3f} tail {kern[3]:.3f}")

    # ---- FP64 plain batched and per-family launches (the comparison forms)
    step(_lib.FP64_PLAIN)
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        step(_lib.FP64_PLAIN)
    fp64_batched_ms = 1e3 * (time.perf_counter() - t) / reps
    per_fam = None
    if not args.no_per_family:
        t = time.perf_counter()
        step(_lib.FP64_PLAIN, launch=_lib.PER_FAMILY)
        pf_s = time.perf_counter() - t
        per_fam = n_fam / pf_s
        log(f"fp64 plain per-family launches: {pf_s:.3f} s, {per_fam:.4g} families/s; "
            f"fp64 plain batched {fp64_batched_ms:.3f} ms")

    if args.copy_probe:
        copy_probe()

    # ---- roofline of the dominant kernel (algorithmic bytes / its duration)
    b_fwd, b_grad = algorithmic_bytes(st, 4, x_fwd=4 * st["clades"])
    fwd_gbs = b_fwd / (kern[1] / 1e3) / 1e9
    bwd_gbs = (b_grad - b_fwd) / (kern[2] / 1e3) / 1e9
    if kern[2] >= kern[1]:
        dom, ach, traffic_bytes, dom_ms = "k_backward", bwd_gbs, b_grad - b_fwd, kern[2]
    else:
        dom, ach, traffic_bytes, dom_ms = "k_forward", fwd_gbs, b_fwd, kern[1]
    log(f"roofline: fwd {fwd_gbs:.1f} GB/s ({b_fwd / 1e9:.3f} GB), bwd {bwd_gbs:.1f} GB/s "
        f"({(b_grad - b_fwd) / 1e9:.3f} GB), whole step {b_grad / (ms_step / 1e3) / 1e9:.1f} GB/s")
    hbm = measured_traffic(dom) if args.workload == "config4" and args.families is None else None

    # ---- the sampler-style loop (config5), its own workload
    samp = None
    if not args.no_sampler:
        samp = sampler(w, synth, n_points=args.sampler_points, tune_cache=args.tune_cache or None,
                       tune=not args.no_tune)

    # ---- CPU baseline: the oracle (C++ restatement) on a bounded sample
    cpu = cpu2 = calib = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, n_fam, cpu_threads(), args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
        # the same families' share of work at 2 threads: a sample as many times smaller
        cpu2 = cpu_baseline(wl, n_fam, 2, args.cpu_seconds, k=cpu["every"] * max(1, cpu["cores"] // 2))
        log(f"cpu baseline at 2 threads: {cpu2}")
        calib = cpu_calibration()
        log(f"cpu calibration: {calib}")

    out = {
        "metric": "families/s for logL+grad: FP32-scaled batched vs FP64 plain, rel err ≤1e-6",
        "value": value,
        "unit": "families/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n_fam} families, 50-taxon tree + 4 WGDs, dt=0.01, "
                               "logL+gradient, FP32 with per-clade exponents, one batched launch",
                   "families": n_fam, "cells": st["cells"], "split_terms": st["split_terms"]},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": hbm, "kernel": dom,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, scaled by the "
                                        "counters' reading of a 256 MiB copy, committed in profiles/"
                                        "config4_fp32_gradient_hbm_traffic.md (a profiler run of its own, not "
                                        "this process)") if hbm else None,
                     "algorithmic_bytes": traffic_bytes, "duration_ms": dom_ms,
                     "duration": "HIP events on the library's stream around the kernel's phase (all size-class "
                                 "launches and the FP64 families, several streams) of eager evaluations, median"},
        "cpu_baseline": cpu,
        "cpu_baseline_2_threads": cpu2,
        "cpu_calibration_kat2": calib,
        "accuracy": {"fp32_vs_fp64_family_max_rel": rel_fam, "fp32_vs_fp64_total_rel": rel_tot,
                     "fp32_vs_fp64_grad_rel": gerr, "fp64_plain_underflowed_families": n_underflow,
                     "fp32_families_recomputed_in_fp64": n_fallback},
        "fp64_plain_per_family_families_per_s": per_fam,
        "speedup_vs_fp64_per_family": (value / per_fam) if per_fam else None,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_2_threads": (value / cpu2["value"]) if cpu2 else None,
        "sampler": samp,
        "tuning": tune,
        "fp64_plain_batched_ms": fp64_batched_ms,
        "kernel_ms": {"tables": kern[0], "forward": kern[1], "backward": kern[2], "tail": kern[3]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
