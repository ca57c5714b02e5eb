#!/usr/bin/env python3
"""Profiling driver: runs the operator passes of one workload a fixed number
of times, plus one device-to-device copy of known size, so that a
`rocprofv3 --pmc` run yields per-dispatch counters for each pass and a
calibration point for the counters (the copy moves exactly --copy-mib MiB
in and out).

usage: python3 topo-dda-forwin64_amd/profile_passes.py [--workload topo80|mixed128|batch64]
       [--precision c128|c64] [--reps 20] [--copy-mib 256] [--no-tune] [--plan-out plan.json]
       [--incidences N]

A single operator (topo80, mixed128) runs the autotuner first (as in
bench.py), so its tuner's trial launches precede the timed passes in a
trace; batch64 keeps the planner's tiles, as bench.py's batch does.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import topodda as T  # noqa: E402
import workloads as W  # noqa: E402


def copy_calibration(mib):
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="topo80")
    ap.add_argument("--precision", default="c128")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true", help="planner tiles instead of the autotuner's")
    ap.add_argument("--incidences", type=int, default=0,
                    help="batch64: right-hand sides per structure, the first N of the 8 incidences (0: all)")
    ap.add_argument("--plan-out", default=None,
                    help="JSON file: per precision, whether each operator runs fused z + y passes "
                         "(profile_summary.py traffic reads it from the parent of the counter directories)")
    a = ap.parse_args()
    ctx = T.Context(0)
    prec = T.C128 if a.precision == "c128" else T.C64
    if a.workload == "batch64":
        # the headline batch as bench.py builds it: planner tiles, 8
        # incidences per structure, one launch per FFT length and pass
        probs, incs = W.batch64(seed=4), W.incidences()
        if a.incidences > 0:
            incs = incs[:a.incidences]
        cores, models = [], []
        for p in probs:
            core = T.AProductCore(ctx, p.R, p.Nx, p.Ny, p.Nz, p.d, p.lam)
            m = T.DeviceModel(core, len(incs))
            rhs = [p.rhs(nK, nE)[:2] for nK, nE in incs]
            m.set(T.VEC_AL, np.concatenate([al for al, _ in rhs]))
            m.set(T.VEC_P, np.concatenate([e for _, e in rhs]))
            cores.append(core)
            models.append(m)
        b = T.Batch(models)
        ms = b.time_passes(prec, reps=a.reps)
        grid = sorted({c.grid for c in cores})
    else:
        if a.workload == "topo80":
            prob, _ = W.topo80()
        elif a.workload == "mixed128":
            prob = W.mixed128()
        else:
            raise SystemExit("unknown workload")
        al, E, _ = prob.rhs(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
        core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam)
        if not a.no_tune:
            core.autotune()  # the tiles and x-pass kernel bench.py measures
        m = T.DeviceModel(core, 1)
        m.set(T.VEC_AL, al)
        m.set(T.VEC_P, E)
        ms = m.time_passes(prec, reps=a.reps)
        grid = core.grid
    if a.plan_out:
        cl = cores if a.workload == "batch64" else [core]
        with open(a.plan_out, "w") as fh:
            plan = {p: [c.zy_fused(q) for c in cl] for p, q in (("c128", T.C128), ("c64", T.C64))}
            plan["grids"] = [list(c.grid) for c in cl]
            plan["nrhs"] = len(incs) if a.workload == "batch64" else 1
            json.dump(plan, fh)
    copy_calibration(a.copy_mib)
    print(json.dumps({"workload": a.workload, "precision": a.precision, "grid": grid,
                      "pass_ms": [float(x) for x in ms], "copy_mib": a.copy_mib}))


if __name__ == "__main__":
    main()
