#!/usr/bin/env python3
"""Summarises rocprofv3 CSV output of the operator passes.

  python3 topo-dda-forwin64_amd/profile_summary.py stats  <kernel_stats.csv>  [--top 20]
  python3 topo-dda-forwin64_amd/profile_summary.py last <kernel_trace.csv> --kernel 'k_pxr<double, 160>' [--last 200]
  python3 topo-dda-forwin64_amd/profile_summary.py span <kernel_trace.csv> --precision c64 --workload batch64 [--last 10]
  python3 topo-dda-forwin64_amd/profile_summary.py window <kernel_trace.csv> --kernel 'k_pxr<float' --reps 20 --per 13 [--skip 0] [--distinct]
  python3 topo-dda-forwin64_amd/profile_summary.py traffic <fetch counter_collection.csv>
          <write counter_collection.csv> --precision c128|c64 --workload topo80|mixed128|batch64
          [--json profiles/traffic.json]

`stats` prints the per-kernel duration table (markdown).  `last` prints the
mean, min and max duration of the last N dispatches of one kernel (the
back-to-back timing launches that end a bench run).  `span` prints, per
pass, the wall span of each of the last N repetitions of the pass (first
launch start to last launch end; what HIP events around the pass measure)
and the summed kernel time.  `window` takes the dispatches whose name
contains --kernel, drops the last --skip of them, splits the last
--reps x --per of the rest into --reps runs of --per consecutive dispatches
(one repetition of a pass over a batch: one launch per FFT length), and
prints the wall span of each run (what HIP events around it measure).
`traffic` prints the
average FETCH_SIZE / WRITE_SIZE per dispatch of each pass kernel, scaled by
the calibration that the plain device copy in the same run gives (the copy
of --copy-mib MiB run by profile_passes.py reads and writes exactly that
many bytes), next to the algorithmic bytes of the pass; with --json it
merges {"<workload>_<precision>:<pass name>": measured HBM bytes per pass}
into that file (bench.py reads profiles/traffic.json).
"""
import argparse
import collections
import csv
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

# pass kernels: name patterns (the x and y passes have register-resident and
# LDS-staged kernels; the first pattern with dispatches wins) and the pass
# names of bench.py
PASS_NAMES = {("k_p1<",): "k_p1_zfwd", ("k_pyr<{t}, false", "k_py<{t}, false"): "k_py_yfwd",
              ("k_pxr<", "k_px<"): "k_px_xgreen", ("k_pyr<{t}, true", "k_py<{t}, true"): "k_py_yinv",
              ("k_p5<",): "k_p5_zinv"}
# operators with fused z + y passes run k_pzy in the P2 slot (P1 + P2) and
# k_pyz in the P5 slot (P4 + P5), beside the separate passes of the others
FUSED = {"k_py_yfwd": "k_pzy<{t}", "k_p5_zinv": "k_pyz<{t}"}


def pass_kernels(names, pats, pname, t, contains=False):
    """Kernel names of one pass slot: those of the first pattern that has
    dispatches, plus the slot's fused kernels."""
    def match(kn, pat):
        full = f"tdda::{pat.format(t=t)}"
        return (full in kn if contains else kn.startswith(full)) and f"<{t}" in kn
    sel = set()
    for pat in pats:
        sel = {kn for kn in names if match(kn, pat)}
        if sel:
            break
    if pname in FUSED:
        sel |= {kn for kn in names if match(kn, FUSED[pname])}
    return sel


def stats(path, top):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | avg us | min us | max us | % time |")
    print("|---|---|---|---|---|---|")
    for r in rows[:top]:
        name = r["Name"].split("(")[0].replace("void ", "")
        print(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
              f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")


def last(path, kernel, n):
    """Mean, min and max of the last n dispatches of every kernel whose name
    starts with `kernel` (a batch runs one kernel per FFT length); with
    several, also the sum of their means (the duration of one pass over the
    batch, launch gaps excluded)."""
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        kn = r["Kernel_Name"].replace("void ", "")
        if kn.startswith("tdda::" + kernel):
            per[kn.split("(")[0]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = 0.0
    for kn, d in sorted(per.items()):
        tail = d[-n:]
        total += sum(tail) / len(tail)
        print(f"`{kn}`: {len(d)} dispatches; last {len(tail)}: mean {sum(tail) / len(tail):.2f} us, "
              f"min {min(tail):.2f} us, max {max(tail):.2f} us")
    if len(per) > 1:
        print(f"sum of the means of {len(per)} kernels: {total:.2f} us")


def span(path, precision, workload, reps):
    """Duration of one pass as a HIP-event pair around it sees it: the pass's
    dispatches (a batch runs one launch per FFT length, spread over several
    streams) are split into the last `reps` repetitions in start order, and
    each repetition spans from its first start to its last end.  Also the sum
    of the repetition's kernel durations (busy time, overlap counted twice)."""
    t = "double" if precision == "c128" else "float"
    rows = []
    for r in csv.DictReader(open(path)):
        kn = r["Kernel_Name"].replace("void ", "").split("(")[0]
        rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kn))
    rows.sort()
    print("| pass | launches per pass | repetitions | span us (mean) | span us (min) | kernel busy us (mean) |")
    print("|---|---|---|---|---|---|")
    for pats, pname in PASS_NAMES.items():
        names = pass_kernels({kn for _, _, kn in rows}, pats, pname, t)
        if not names:
            continue
        d = [r for r in rows if r[2] in names]
        if workload != "batch64":
            d = [r for r in d if r[2] == d[-1][2]]
            g = 1
        else:
            g = len(names)
        d = d[-reps * g:]
        spans, busy = [], []
        for k in range(0, len(d) - g + 1, g):
            blk = d[k:k + g]
            spans.append((max(e for _, e, _ in blk) - min(s for s, _, _ in blk)) / 1e3)
            busy.append(sum(e - s for s, e, _ in blk) / 1e3)
        print(f"| {pname} | {g} | {len(spans)} | {sum(spans) / len(spans):.2f} | {min(spans):.2f} | "
              f"{sum(busy) / len(busy):.2f} |")


def sq(counter_csvs, trace_csv, top, clock_ghz=2.4, simds=1024):
    """Per-kernel SQ counter summary of one or more --pmc runs and a kernel
    trace of the same workload (durations): VALU busy = VALU instructions x
    4 cycles (a wave64 instruction on a 16-lane SIMD) over the SIMD cycles
    of the kernel's mean duration; wait = SQ_WAIT_ANY / SQ_WAVE_CYCLES; LDS
    bank-conflict cycles per LDS-active cycle; instructions per wave."""
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in counter_csvs:
        for r in csv.DictReader(open(f)):
            kn = r["Kernel_Name"].split("(")[0].replace("void ", "")
            agg[kn][r["Counter_Name"]].append(float(r["Counter_Value"]))
    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):
        kn = r["Kernel_Name"].split("(")[0].replace("void ", "")
        dur[kn].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    print("| kernel | us | VALU busy % | wait % of wave cycles | LDS bank-conflict cycles per LDS-active cycle "
          "| per wave: VALU, LDS, SALU, VMEM rd, VMEM wr |")
    print("|---|---|---|---|---|---|")
    names = [kn for kn in agg if kn in dur and any(k in kn for k in ("k_pxr", "k_pzy", "k_pyz", "k_pyr", "k_p1", "k_p5"))]
    for kn in sorted(names, key=lambda k: -sum(dur[k]))[:top]:
        c = {n: sum(v) / len(v) for n, v in agg[kn].items()}
        us = sum(dur[kn]) / len(dur[kn])
        cyc = us * 1e-6 * clock_ghz * 1e9 * simds
        w = c.get("SQ_WAVES", 1) or 1
        print(f"| {kn} | {us:.1f} | {100 * c.get('SQ_INSTS_VALU', 0) * 4 / cyc:.0f} | "
              f"{100 * c.get('SQ_WAIT_ANY', 0) / max(c.get('SQ_WAVE_CYCLES', 1), 1):.0f} | "
              f"{c.get('SQ_LDS_BANK_CONFLICT', 0) / max(c.get('SQ_ACTIVE_INST_LDS', 1), 1):.2f} | "
              f"{c.get('SQ_INSTS_VALU', 0) / w:.0f}, {c.get('SQ_INSTS_LDS', 0) / w:.0f}, "
              f"{c.get('SQ_INSTS_SALU', 0) / w:.0f}, {c.get('SQ_INSTS_VMEM_RD', 0) / w:.0f}, "
              f"{c.get('SQ_INSTS_VMEM_WR', 0) / w:.0f} |")


LAST_ROW = {}


def per_kernel(path, counter):
    """Per-dispatch counter values of each kernel, in dispatch order; the row
    of each kernel's last dispatch goes to LAST_ROW."""
    agg = collections.defaultdict(list)
    for k, r in enumerate(csv.DictReader(open(path))):
        if r["Counter_Name"] == counter:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)
            LAST_ROW[(counter, r["Kernel_Name"])] = k
    return agg


def traffic(fetch_csv, write_csv, precision, workload, copy_mib, out_json, last_n=0):
    """Per-dispatch HBM bytes of each pass; last_n > 0 averages only the last
    last_n dispatches of each pass kernel (those of profile_passes.py's
    timing, after the autotuner's trial launches)."""
    import workloads as W
    t = "double" if precision == "c128" else "float"
    s = 16 if precision == "c128" else 8
    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    if last_n > 0:
        for d in (fetch, write):
            for kn in d:
                if not kn.startswith("__amd"):
                    d[kn] = d[kn][-last_n:]
    copy_bytes = copy_mib * 1024.0 * 1024.0
    # the copy may be split into several dispatches: take the largest ones
    # that together move copy_bytes of writes
    cw = sorted(write.get("__amd_rocclr_copyBuffer", []), reverse=True)
    cf = sorted(fetch.get("__amd_rocclr_copyBuffer", []), reverse=True)
    acc, n = 0.0, 0
    while n < len(cw) and acc < 0.999 * copy_bytes:
        acc += cw[n]
        n += 1
    wfac = copy_bytes / sum(cw[:n])
    ffac = copy_bytes / sum(cf[:n])
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import batch_bytes, pass_bytes

    # which operators ran fused z + y passes: profile_passes.py's plan file
    # beside the counter files (one flag per structure, in workload order)
    plan, grids, nrhs = None, None, len(W.incidences())
    pf = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(fetch_csv))), "plan.json")
    if os.path.exists(pf):
        with open(pf) as fh:
            pj = json.load(fh)
        plan, grids, nrhs = pj.get(precision), pj.get("grids"), pj.get("nrhs", nrhs)

    class Core:
        def zy_fused(self, prec):
            return self.fused

    def core_of(p, k=0):
        c = Core()
        c.Nx, c.Ny, c.Nz, c.N = p.Nx, p.Ny, p.Nz, p.N
        # the operator's grid as the planner chose it (plan file), else the
        # smallest smooth lengths
        c.grid = tuple(grids[k]) if grids else tuple(next_smooth(2 * n - 1) for n in (p.Nx, p.Ny, p.Nz))
        c.fused = bool(plan[k]) if plan else False
        return c
    if workload == "batch64":
        alg = batch_bytes([(core_of(p, k), nrhs) for k, p in enumerate(W.batch64(seed=4))], s)
    else:
        alg = pass_bytes(core_of(W.topo80()[0] if workload == "topo80" else W.mixed128()), s)
    print(f"calibration: device copy of {copy_mib} MiB in {n} dispatches; FETCH_SIZE x {ffac:.3f}, "
          f"WRITE_SIZE x {wfac:.3f}")
    print("| pass | kernels | dispatches | fetch MB | write MB | measured MB | algorithmic MB | ratio |")
    print("|---|---|---|---|---|---|---|---|")
    res = {}
    for k, (pats, pname) in enumerate(PASS_NAMES.items()):
        # the first pattern with dispatches; a batch has one kernel per FFT
        # length, whose per-dispatch bytes add up to one pass over the batch
        names = sorted(pass_kernels(set(fetch), pats, pname, t, contains=True))
        if not names:
            continue
        if workload != "batch64":
            # one kernel per pass: the last one dispatched (the autotuner's
            # trials of other kernels come first)
            names = [max(names, key=lambda kn: LAST_ROW[("FETCH_SIZE", kn)])]
        f = sum(sum(fetch[kn]) / len(fetch[kn]) for kn in names) * ffac
        w = sum(sum(write.get(kn, [0.0])) / max(1, len(write.get(kn, []))) for kn in names) * wfac
        res[f"{workload}_{precision}:{pname}"] = int(f + w)
        label = names[0].split('(')[0].replace('void ', '') if len(names) == 1 else f"{len(names)} kernels"
        print(f"| {pname} | `{label}` | {sum(len(fetch[kn]) for kn in names)} | {f / 1e6:.2f} | "
              f"{w / 1e6:.2f} | {(f + w) / 1e6:.2f} | {alg[k] / 1e6:.2f} | {(f + w) / alg[k]:.2f} |")
    if out_json:
        # merged into the file's other workloads
        prev = {}
        if os.path.exists(out_json):
            with open(out_json) as fh:
                prev = json.load(fh)
        prev.update(res)
        with open(out_json, "w") as fh:
            json.dump(prev, fh, indent=1, sort_keys=True)


def window(trace_csv, kernel, reps, per, skip, distinct=False):
    """Wall span of each of the last `reps` runs of `per` consecutive
    dispatches of the kernels whose name contains `kernel`.  With
    `distinct`, the runs end where the last `reps` x `per` dispatches in
    which every run holds the same `per` different kernels end (one launch
    per FFT length of a batch, as opposed to later single-operator
    launches), and `skip` counts back from there."""
    with open(trace_csv) as fh:
        rows = [r for r in csv.DictReader(fh) if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    need = reps * per
    end = len(rows)
    if distinct:
        def ok(e):
            # every run: the same `per` different kernels
            if e < need:
                return False
            first = {r["Kernel_Name"] for r in rows[e - per:e]}
            return len(first) == per and all({r["Kernel_Name"] for r in rows[e - (k + 1) * per:e - k * per]} == first
                                             for k in range(1, reps))
        while end >= need and not ok(end):
            end -= 1
        if end < need:
            sys.exit(f"no {reps} runs of {per} different {kernel!r} kernels")
    rows = rows[:end - skip] if skip else rows[:end]
    if len(rows) < need:
        sys.exit(f"{len(rows)} dispatches match {kernel!r}, {need} needed")
    rows = rows[-need:]
    spans, busy = [], []
    for k in range(reps):
        run = rows[k * per:(k + 1) * per]
        s0 = min(int(r["Start_Timestamp"]) for r in run)
        e1 = max(int(r["End_Timestamp"]) for r in run)
        spans.append((e1 - s0) / 1e3)
        busy.append(sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in run) / 1e3)
    names = sorted({r["Kernel_Name"].split("(")[0].replace("void ", "") for r in rows})
    whole = (max(int(r["End_Timestamp"]) for r in rows) - int(rows[0]["Start_Timestamp"])) / 1e3
    print(f"`{kernel}`: {reps} runs of {per} dispatches ({len(names)} kernels), skipping {skip}: "
          f"span mean {sum(spans) / reps:.2f} us, min {min(spans):.2f} us, max {max(spans):.2f} us; "
          f"first start to last end / runs {whole / reps:.2f} us; kernel busy (summed) mean {sum(busy) / reps:.2f} us")
    return spans


def next_smooth(n):
    m = n
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "traffic", "last", "span", "sq", "window"])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--precision", default="c128")
    ap.add_argument("--workload", default="topo80")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--json", default=None)
    ap.add_argument("--kernel", default="k_pxr<double, 160>")
    ap.add_argument("--last", type=int, default=200,
                    help="last: dispatches averaged; traffic: last dispatches per pass kernel (0: all)")
    ap.add_argument("--reps", type=int, default=20, help="window: runs")
    ap.add_argument("--per", type=int, default=1, help="window: dispatches per run")
    ap.add_argument("--skip", type=int, default=0, help="window: matching dispatches dropped from the end")
    ap.add_argument("--distinct", action="store_true",
                    help="window: end at the last runs that each hold --per different kernels")
    a = ap.parse_args()
    if a.mode == "window":
        window(a.files[0], a.kernel, a.reps, a.per, a.skip, a.distinct)
    elif a.mode == "stats":
        stats(a.files[0], a.top)
    elif a.mode == "last":
        last(a.files[0], a.kernel, a.last)
    elif a.mode == "span":
        span(a.files[0], a.precision, a.workload, a.last)
    elif a.mode == "sq":
        # files: the counter CSVs, then the kernel-trace CSV
        sq(a.files[:-1], a.files[-1], a.top)
    else:
        traffic(a.files[0], a.files[1], a.precision, a.workload, a.copy_mib, a.json, a.last)


if __name__ == "__main__":
    main()
