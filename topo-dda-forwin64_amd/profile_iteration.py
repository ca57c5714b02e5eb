#!/usr/bin/env python3
"""Anatomy of one BiCGSTAB iteration of the batched mixed-precision solve
(64 structures x 8 incidences, complex64 inner solves, planner tiles as in
bench.py).

usage:
  python3 topo-dda-forwin64_amd/profile_iteration.py sweep
      marginal cost of an inner iteration against the number of right-hand
      sides still iterating: one refinement cycle (max_outer = 1) capped at
      inner_max_it = k for increasing k; then the whole solve for a few
      inner tolerances, alternated, medians over the repetitions
  python3 topo-dda-forwin64_amd/profile_iteration.py trace [--its 20]
      two capped solves (max_outer = 1, inner_max_it = its) for
      `rocprofv3 --kernel-trace --output-format csv`
  python3 topo-dda-forwin64_amd/profile_iteration.py spans CSV
      reads that kernel trace: the iteration loop of the second solve (between
      its two k_mix launches), busy and idle time of the device, and per
      operator application the span of each pass kind over its launch groups
"""
import argparse
import collections
import csv
import os
import re
import statistics
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

TOL = 1e-8


def build_batch(T, W, ctx):
    probs, incs = W.batch64(seed=4), W.incidences()
    cores = [T.AProductCore(ctx, p.R, p.Nx, p.Ny, p.Nz, p.d, p.lam, p.nback) for p in probs]
    models = []
    for p, core in zip(probs, cores):
        rhs = [p.rhs(nK, nE)[:2] for nK, nE in incs]
        m = T.DeviceModel(core, len(incs))
        m.set(T.VEC_AL, np.concatenate([a for a, _ in rhs]))
        m.set(T.VEC_E, np.concatenate([e for _, e in rhs]))
        models.append(m)
    return T.Batch(models), cores, models


def timed(ctx, fn):
    ctx.sync()
    t0 = time.perf_counter()
    out = fn()
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3, out


def sweep(T, W, ctx):
    batch, _, _ = build_batch(T, W, ctx)
    for _ in range(2):
        batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
    print("| inner_max_it | ms (max_outer = 1) | right-hand sides still iterating | marginal ms / iteration |")
    print("|---|---|---|---|")
    prev = None
    for k in (1, 10, 20, 30, 34, 38, 42, 46, 50, 55):
        ms = []
        for _ in range(3):
            d, inf = timed(ctx, lambda: batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1,
                                                    max_outer=1, inner_max_it=k))
            ms.append(d)
        d = statistics.median(ms)
        live = sum(1 for i in inf if i["iterations"] >= k)
        marg = f"{(d - prev[1]) / (k - prev[0]):.3f}" if prev else ""
        print(f"| {k} | {d:.2f} | {live} | {marg} |", flush=True)
        prev = (k, d)
    vals = (1e-4, 2e-4, 3e-4, 5e-4)
    res = {v: [] for v in vals}
    info = {}
    for v in vals:
        batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1, inner_tol=v)
    for _ in range(5):
        for v in vals:
            d, info[v] = timed(ctx, lambda: batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1,
                                                        inner_tol=v))
            res[v].append(d)
    print()
    print("| inner_tol | median ms per solve of the batch | inner iterations (mean) | refinement cycles (mean) | "
          "all converged |")
    print("|---|---|---|---|---|")
    for v in vals:
        its = np.mean([i["iterations"] for i in info[v]])
        outer = np.mean([i["outer"] for i in info[v]])
        ok = all(i["status"] == T.CONVERGED and i["true_residual"] <= TOL for i in info[v])
        print(f"| {v:g} | {statistics.median(res[v]):.1f} | {its:.1f} | {outer:.2f} | {ok} |", flush=True)
    batch.close()


def trace(T, W, ctx, its):
    batch, _, _ = build_batch(T, W, ctx)
    for _ in range(2):
        d, _ = timed(ctx, lambda: batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1,
                                              max_outer=1, inner_max_it=its))
        print(f"capped solve ({its} iterations): {d:.2f} ms", flush=True)
    batch.close()


def pass_kind(name):
    for k in ("k_pzy", "k_pyz", "k_pxr", "k_px<", "k_py<", "k_pyr", "k_p1", "k_p5"):
        if k in name:
            return k.rstrip("<")
    return None


def spans(path):
    """Prints the tables; returns {"loop_ms", "busy_ms", "idle_ms", "span_us": {kind: (p, t)}}."""
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    for r in rows:
        r["s"], r["e"] = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
    rows.sort(key=lambda r: r["s"])
    mix = [r for r in rows if "k_mix" in r["Kernel_Name"]]
    if len(mix) < 4:
        sys.exit("expected two solves (two k_mix launches each) in the trace")
    # the second solve's iteration loop: after its demotion k_mix, before its accumulation k_mix
    sel = [r for r in rows if r["s"] > mix[2]["e"] and r["e"] < mix[3]["s"]]
    t0, t1 = sel[0]["s"], max(r["e"] for r in sel)
    busy, cs, ce = 0, None, None
    for r in sel:
        if ce is None or r["s"] > ce:
            if ce is not None:
                busy += ce - cs
            cs, ce = r["s"], r["e"]
        else:
            ce = max(ce, r["e"])
    busy += ce - cs
    fins = [r for r in sel if "k_finish" in r["Kernel_Name"]]
    print(f"iteration loop: {(t1 - t0) / 1e6:.2f} ms, device busy {busy / 1e6:.2f} ms, idle {(t1 - t0 - busy) / 1e6:.2f} "
          f"ms, {len(sel)} dispatches, {len(fins)} finishing launches")
    # operator applications: the pass launches between two finishing launches
    apps, cur = [], []
    for r in sel:
        if "k_finish" in r["Kernel_Name"]:
            if cur:
                apps.append(cur)
            cur = []
        elif pass_kind(r["Kernel_Name"]):
            cur.append(r)
    span = collections.defaultdict(lambda: ([], []))
    for i, a in enumerate(apps):
        g = collections.defaultdict(list)
        for r in a:
            g[pass_kind(r["Kernel_Name"])].append(r)
        for k, rs in g.items():
            span[k][i % 2].append((max(r["e"] for r in rs) - min(r["s"] for r in rs)) / 1e3)
        span["application"][i % 2].append((max(r["e"] for r in a) - min(r["s"] for r in a)) / 1e3)
    print()
    print("| pass kind | launches per application | span us, p / Ap application | span us, t / At application |")
    print("|---|---|---|---|")
    per_app = collections.Counter(pass_kind(r["Kernel_Name"]) for a in apps for r in a)
    for k, (ev, od) in span.items():
        n = per_app[k] / len(apps) if k != "application" else sum(per_app.values()) / len(apps)
        print(f"| {k} | {n:.0f} | {statistics.mean(ev):.1f} | {statistics.mean(od):.1f} |")
    agg = collections.defaultdict(lambda: [0, 0])
    for r in sel:
        n = re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "")
        agg[n][0] += 1
        agg[n][1] += r["e"] - r["s"]
    print()
    print("| kernel | dispatches | total ms (all streams) |")
    print("|---|---|---|")
    for n, (c, d) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:12]:
        print(f"| `{n}` | {c} | {d / 1e6:.3f} |")
    return {"loop_ms": (t1 - t0) / 1e6, "busy_ms": busy / 1e6, "idle_ms": (t1 - t0 - busy) / 1e6,
            "span_us": {k: (statistics.mean(ev), statistics.mean(od)) for k, (ev, od) in span.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["sweep", "trace", "spans"])
    ap.add_argument("csv", nargs="?")
    ap.add_argument("--its", type=int, default=20)
    a = ap.parse_args()
    if a.mode == "spans":
        if not a.csv:
            sys.exit("spans needs the kernel trace CSV")
        spans(a.csv)
        return
    import topodda as T
    import workloads as W
    ctx = T.Context(0)
    if a.mode == "sweep":
        sweep(T, W, ctx)
    else:
        trace(T, W, ctx, a.its)
    ctx.close()


if __name__ == "__main__":
    main()
