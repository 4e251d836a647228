"""Summaries of rocprofv3 output for the committed profiles.

    python whale.jl_amd/rocprof_summary.py trace DIR
        per-launch kernel durations (grouped by kernel, grid, workgroup and
        LDS size) and, per FP32 logL+gradient evaluation, the time covered by
        any launch of k_forward / k_backward (union over streams), split into
        evaluations whose phases run apart (per-kernel timing, eager) and
        evaluations whose phases overlap (graph replays);
    python whale.jl_amd/rocprof_summary.py pmc DIR
        counters per kernel from a --pmc run: FETCH_SIZE / WRITE_SIZE scaled
        by a 256 MiB device-to-device copy in the same run (bench.py
        --copy-probe) to bytes per evaluation, and SQ_* counters as sums per
        evaluation with the derived busy / wait / VALU / LDS ratios;
    python whale.jl_amd/rocprof_summary.py traffic FETCH_DIR WRITE_DIR
        the two byte counters' runs combined into MB per evaluation.

DIR is the -d directory of the rocprofv3 run, searched recursively for its
result databases (rocpd SQLite, the default output format of rocprofv3).
Output is Markdown on stdout.
"""
import glob
import os
import re
import sqlite3
import statistics
import sys
from collections import defaultdict


def _dbs(d):
    files = sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True))
    if not files:
        raise SystemExit(f"no rocpd database under {d}")
    return files


def _query(d, sql):
    for f in _dbs(d):
        con = sqlite3.connect(f)
        try:
            yield from con.execute(sql)
        finally:
            con.close()


def short_name(n):
    """`k_backward<float,true,true>` from a demangled kernel name."""
    n = n.strip()
    m = re.search(r"(k_\w+|__amd_rocclr_\w+)(<[^()]*>)?", n)
    if not m:
        return n[:60]
    return m.group(1) + (m.group(2) or "").replace(" ", "").replace("whale::", "")


def _launches(d):
    out = []
    for name, start, end, gx, wx, lds in _query(d, "select name, start, end, grid_x, workgroup_x, lds_size "
                                                   "from kernels"):
        out.append(dict(name=short_name(name), start=int(start), end=int(end), wg=int(wx),
                        grid=int(gx) // max(1, int(wx)), lds=int(lds)))
    out.sort(key=lambda x: x["start"])
    return out


def _evaluations(launches):
    """Launches in start order split into evaluations: each ends with its
    k_tail (the batch sum), and holds what was launched since the previous
    one (its k_tables when parameters were set, its fills and DP kernels)."""
    evs, cur = [], []
    for x in launches:
        cur.append(x)
        if x["name"].startswith("k_tail"):
            evs.append(cur)
            cur = []
    return evs


def _union(iv):
    iv = sorted(iv)
    tot, cs, ce = 0, None, None
    for s, e in iv:
        if cs is None or s > ce:
            if cs is not None:
                tot += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    if cs is not None:
        tot += ce - cs
    return tot


def _stats(xs):
    return statistics.median(xs), sum(xs) / len(xs), min(xs), max(xs)


def trace(d):
    L = _launches(d)
    print("## Per launch (ms): kernel, grid (workgroups), workgroup size, LDS bytes\n")
    print("| kernel | grid | wg | LDS | calls | avg | min | max |")
    print("|---|---|---|---|---|---|---|---|")
    g = defaultdict(list)
    for x in L:
        g[(x["name"], x["grid"], x["wg"], x["lds"])].append((x["end"] - x["start"]) / 1e6)
    for k, v in sorted(g.items(), key=lambda kv: -sum(kv[1]) / len(kv[1])):
        print(f"| `{k[0]}` | {k[1]} | {k[2]} | {k[3]} | {len(v)} | {sum(v) / len(v):.3f} | {min(v):.3f} | "
              f"{max(v):.3f} |")
    evs = _evaluations(L)
    res = {"apart": defaultdict(list), "overlapped": defaultdict(list)}
    for ev in evs:
        fwd = [x for x in ev if x["name"].startswith("k_forward")]
        bwd = [x for x in ev if x["name"].startswith("k_backward")]
        f32 = [x for x in fwd if "<float" in x["name"]]
        b32 = [x for x in bwd if "<float" in x["name"]]
        if not b32 or not f32:
            continue  # FP32 logL+gradient evaluations only
        kind = "apart" if max(x["end"] for x in f32) <= min(x["start"] for x in b32) else "overlapped"
        r = res[kind]
        if kind == "apart":
            # the reverse phase: from the first FP32 reverse sweep on (the FP64
            # evaluation of routed families runs beside the forward phase)
            t0 = min(x["start"] for x in b32)
            r["k_forward (all classes + FP64 families), up to the reverse phase"].append(
                _union([(x["start"], min(x["end"], t0)) for x in fwd if x["start"] < t0]) / 1e6)
            r["k_backward (all classes + FP64 families), reverse phase"].append(
                _union([(max(x["start"], t0), x["end"]) for x in bwd if x["end"] > t0]) / 1e6)
        else:
            r["k_forward (all classes + FP64 families)"].append(_union([(x["start"], x["end"]) for x in fwd]) / 1e6)
            r["k_backward (all classes + FP64 families)"].append(_union([(x["start"], x["end"]) for x in bwd]) / 1e6)
        r["first to last kernel"].append((max(x["end"] for x in ev) - min(x["start"] for x in ev)) / 1e6)
    print("\n## Per FP32 logL+gradient evaluation (ms): time covered by any launch of the kernel "
          "(union over streams)")
    titles = {"apart": "phases apart (per-kernel timing: all forwards end before the first reverse sweep)",
              "overlapped": "phases overlapped (graph replays: each stream's reverse sweeps follow its forwards)"}
    for kind in ("apart", "overlapped"):
        r = res[kind]
        if not r:
            continue
        n = len(next(iter(r.values())))
        print(f"\n{titles[kind]}: {n} evaluations\n")
        print("| | median | mean | min | max |")
        print("|---|---|---|---|---|")
        for k, v in r.items():
            print("| {} | {:.3f} | {:.3f} | {:.3f} | {:.3f} |".format(k, *_stats(v)))


def pmc(d):
    # dispatches in time order, grouped into evaluations (one k_tables launch
    # starts each); only FP32 logL+gradient evaluations are summed
    disp = {}
    for name, di, cname, val, start in _query(d, "select kernel_name, dispatch_id, counter_name, value, start "
                                                 "from counters_collection"):
        x = disp.setdefault(di, dict(name=short_name(name), start=int(start), c=defaultdict(float)))
        x["c"][cname] += float(val)
    order = sorted(disp.values(), key=lambda x: x["start"])
    evs = _evaluations(order)
    fp32 = [ev for ev in evs if any(x["name"].startswith("k_backward<float") for x in ev)]
    per = defaultdict(lambda: defaultdict(float))  # kernel -> counter -> sum over FP32 evaluations
    calls = defaultdict(int)
    for ev in fp32:
        for x in ev:
            calls[x["name"]] += 1
            for c, v in x["c"].items():
                per[x["name"]][c] += v
    copy = defaultdict(float)
    for x in order:
        if x["name"].startswith("__amd_rocclr_copyBuffer"):
            for c, v in x["c"].items():
                copy[c] += v
    counters = sorted({c for x in order for c in x["c"]})
    n_eval = max(1, len(fp32))
    print(f"FP32 logL+gradient evaluations in the run: {n_eval} (of {len(evs)} evaluations); "
          "figures are per FP32 evaluation, summed over the kernel's launches in it\n")
    if any(c.startswith(("FETCH_SIZE", "WRITE_SIZE")) for c in counters):
        scale = {}
        for c in counters:
            ref = copy.get(c, 0.0)
            scale[c] = (256 * 1024 * 1024) / (ref * 1024) if ref else None
            print(f"{c}: the 256 MiB copy reads {ref:.1f} KiB: scale {scale[c] and round(scale[c], 3)}")
        print("\n| kernel | launches per evaluation | " + " | ".join(f"{c} (MB)" for c in counters) + " |")
        print("|---|---|" + "---|" * len(counters))
        for k in sorted(per):
            vals = [per[k][c] * 1024 * (scale[c] or 0) / n_eval / 1e6 for c in counters]
            print(f"| `{k}` | {calls[k] / n_eval:g} | " + " | ".join(f"{v:.1f}" for v in vals) + " |")
    sq = [c for c in counters if c.startswith("SQ_")]
    if sq:
        print("\n| kernel | launches per evaluation | " + " | ".join(sq) + " |")
        print("|---|---|" + "---|" * len(sq))
        for k in sorted(per):
            print(f"| `{k}` | {calls[k] / n_eval:g} | " + " | ".join(f"{per[k][c] / n_eval:.4g}" for c in sq) + " |")
        print("\nDerived, per kernel (ratios of its sums):\n")
        cols = [("wait any / wave cycles", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
                ("VALU active / wave cycles", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
                ("LDS active / wave cycles", "SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
                ("VALU insts per wave", "SQ_INSTS_VALU", "SQ_WAVES"),
                ("LDS insts per wave", "SQ_INSTS_LDS", "SQ_WAVES"),
                ("LDS bank conflict / LDS active", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"),
                ("LDS address conflict / LDS active", "SQ_LDS_ADDR_CONFLICT", "SQ_LDS_IDX_ACTIVE")]
        cols = [t for t in cols if t[1] in sq and t[2] in sq]
        print("| kernel | " + " | ".join(t[0] for t in cols) + " |")
        print("|---|" + "---|" * len(cols))
        for k in sorted(per):
            c = per[k]
            print(f"| `{k}` | " + " | ".join(f"{c[a] / c[b]:.3f}" if c[b] else "–" for _, a, b in cols) + " |")


def _fp32_sums(d):
    """Per kernel: counter sums over the FP32 logL+gradient evaluations of a
    --pmc run, the number of such evaluations, and the copy's sums."""
    disp = {}
    for name, di, cname, val, start in _query(d, "select kernel_name, dispatch_id, counter_name, value, start "
                                                 "from counters_collection"):
        x = disp.setdefault(di, dict(name=short_name(name), start=int(start), c=defaultdict(float)))
        x["c"][cname] += float(val)
    order = sorted(disp.values(), key=lambda x: x["start"])
    evs = _evaluations(order)
    fp32 = [ev for ev in evs if any(x["name"].startswith("k_backward<float") for x in ev)]
    per, calls, copy = defaultdict(lambda: defaultdict(float)), defaultdict(int), defaultdict(float)
    for ev in fp32:
        for x in ev:
            calls[x["name"]] += 1
            for c, v in x["c"].items():
                per[x["name"]][c] += v
    for x in order:
        if x["name"].startswith("__amd_rocclr_copyBuffer"):
            for c, v in x["c"].items():
                copy[c] += v
    return per, calls, max(1, len(fp32)), copy


def traffic(fetch_dir, write_dir):
    """FETCH_SIZE and WRITE_SIZE runs combined: MB per FP32 evaluation and kernel."""
    pf, cf, nf, copy_f = _fp32_sums(fetch_dir)
    pw, cw, nw, copy_w = _fp32_sums(write_dir)
    sf = 256 * 1024 * 1024 / (copy_f["FETCH_SIZE"] * 1024)
    sw = 256 * 1024 * 1024 / (copy_w["WRITE_SIZE"] * 1024)
    print(f"Calibration: the 256 MiB device-to-device copy reads FETCH_SIZE {copy_f['FETCH_SIZE']:.1f} KiB "
          f"and WRITE_SIZE {copy_w['WRITE_SIZE']:.1f} KiB: scales {sf:.3f} and {sw:.3f} (bytes per counted byte).")
    print(f"FP32 logL+gradient evaluations: {nf} in the FETCH_SIZE run, {nw} in the WRITE_SIZE run.\n")
    print("| kernel | launches per evaluation | FETCH (MB) | WRITE (MB) | total (MB) |")
    print("|---|---|---|---|---|")
    for k in sorted(set(pf) | set(pw)):
        if k.startswith("__amd"):
            continue
        f = pf[k]["FETCH_SIZE"] * 1024 * sf / nf / 1e6
        w_ = pw[k]["WRITE_SIZE"] * 1024 * sw / nw / 1e6
        print(f"| `{k}` | {cf[k] / nf:g} | {f:.1f} | {w_:.1f} | {f + w_:.1f} |")


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3])
    elif len(sys.argv) == 3 and sys.argv[1] in ("trace", "pmc"):
        {"trace": trace, "pmc": pmc}[sys.argv[1]](sys.argv[2])
    else:
        raise SystemExit(__doc__)
