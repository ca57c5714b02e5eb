#!/usr/bin/env python3
"""Benchmark: converged DDA solves/s, batched mixed precision against the
plain FP64 device solver, and the operator's HBM roofline.

Headline workload "batch64": 64 structures (Nx, Ny uniform in 16..64, Nz in
8..16; d = 10 nm, lambda = 500 nm; design values U[0, 1] per site, eps
between 1.01 and 2.25 + 0.01i), each with 2 polarisations x 4 incident
angles (0, 10, 20, 30 degrees in the xz plane): 512 right-hand sides with
per-incidence polarizabilities.  One step = one batched solve of all 512:
complex64 inner BiCGSTAB (DDAModel.cpp:131-260 in complex64) inside FP64
refinement until every true residual ||E - A P|| / ||E|| is <= 1e-8,
checked for every right-hand side of every timed step.  `value` is the
converged solves/s of that step.

Beside it:
  * plain FP64: the reference's BiCGSTAB in complex128 on the device, one
    right-hand side at a time (the comparison point), and the batched FP64
    solver;
  * roofline: the fused x pass (FFT, Green product, inverse FFT) of the
    batch in complex64, and of the 128 x 128 x 32 workload in both
    precisions, timed with HIP events in sequence with the other passes
    and alone;
  * field error of every workload against CPU oracle solves converged to
    1e-11, and the oracle's CPU timings on all host threads and on one;
  * the other workloads: single16 (16^3 FP64 solve), topo80 (forward +
    adjoint of a topology step), mixed128 (524k dipoles), serve (Poisson
    request stream), design and periodic (device set-up paths).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       [--workloads single16,topo80,mixed128,design,periodic,serve]
       [--serve-seconds S] [--no-tune] [--batch-tiles planner|tuned]
"""
from __future__ import annotations

import argparse
import collections
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "topo-dda-forwin64_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
TOL = 1e-8
REF_TOL = 1e-11
TUNE = True  # operators run the tile autotuner (--no-tune: planner tiles)
METRIC = "converged DDA solves/s + p50/p99 latency vs plain FP64; Aproduct HBM GB/s"
PASS_NAMES = ["k_p1_zfwd", "k_py_yfwd", "k_px_xgreen", "k_py_yinv", "k_p5_zinv"]


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# algorithmic bytes
# ---------------------------------------------------------------------------
def pass_bytes(core, s):
    """Algorithmic HBM bytes of the five passes of one application
    Y = A P + al o P (one right-hand side), element size s bytes.  An
    operator with fused z + y passes has no P1 / P4 launches: its P2 slot
    holds P1 + P2 (read P, write w2) and its P5 slot P4 + P5 (read w2, al,
    P; write Y) -- w1 never reaches HBM."""
    Nx, Ny, Nz = core.Nx, core.Ny, core.Nz
    Mx, My, Mz = core.grid
    n3 = 3 * core.N
    w1 = Nx * Ny * Mz * 3
    w2 = Nx * My * Mz * 3
    green = (Mx // 2 + 1) * (My // 2 + 1) * (Mz // 2 + 1) * 6
    if core.zy_fused(core_prec(s)):
        return np.array([0, n3 + w2, 2 * w2 + green, 0, w2 + 3 * n3], dtype=np.float64) * s
    return np.array([n3 + w1,              # P1: read P, write w1
                     w1 + w2,              # P2: read w1, write w2
                     2 * w2 + green,       # P3: read/write w2, read Green octant
                     w2 + w1,              # P4: read w2, write w1
                     w1 + 3 * n3],         # P5: read w1, al, P; write Y
                    dtype=np.float64) * s


def core_prec(s):
    import topodda as T
    return T.C128 if s == 16 else T.C64


def batch_bytes(cores_rhs, s):
    """Algorithmic bytes of one application over a batch: every right-hand
    side streams its vectors and intermediates, each structure's Green
    octant is read once for all its right-hand sides."""
    tot = np.zeros(5)
    for core, R in cores_rhs:
        b = pass_bytes(core, s)
        Mx, My, Mz = core.grid
        green = (Mx // 2 + 1) * (My // 2 + 1) * (Mz // 2 + 1) * 6 * s
        b = b * R
        b[2] -= (R - 1) * green
        tot += b
    return tot


def pass_report(ms, b):
    ms = np.asarray(ms, dtype=np.float64)
    return {"ms": [round(float(x), 5) for x in ms],
            "gbs": [round(float(bb / (t * 1e-3) / 1e9), 1) if bb > 0 else None for bb, t in zip(b, ms)],
            "op_gbs": round(float(b.sum() / (ms.sum() * 1e-3) / 1e9), 1),
            "op_frac": round(float(b.sum() / (ms.sum() * 1e-3) / 1e9 / HBM_PEAK_GBS), 4)}


def traffic(key):
    """Measured HBM bytes per dispatch of a pass (rocprofv3 FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy's counters), from profiles/."""
    f = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh).get(key)


def roofline(kernel, device_kernel, algo_bytes, seq_ms, alone_ms, traffic_key, note):
    """HBM roofline of one pass: achieved = algorithmic bytes / duration, the
    duration measured in sequence with the other passes (HIP events around
    the pass inside full applications); the pass launched alone back to
    back beside it."""
    return {"bound": "hbm", "kernel": kernel, "device_kernel": device_kernel,
            "achieved": round(float(algo_bytes / (seq_ms * 1e-3) / 1e9), 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(float(algo_bytes / (seq_ms * 1e-3) / 1e9 / HBM_PEAK_GBS), 4),
            "traffic": traffic(traffic_key), "algorithmic_bytes": int(algo_bytes),
            "duration_us": round(seq_ms * 1e3, 2), "duration_alone_us": round(alone_ms * 1e3, 2),
            "achieved_alone": round(float(algo_bytes / (alone_ms * 1e-3) / 1e9), 1), "note": note}


def converged(infos):
    return all(i["status"] == 1 and i["true_residual"] <= TOL for i in infos)


def lat_stats(ts_s):
    a = np.asarray(ts_s) * 1e3
    return {"p50": round(float(np.percentile(a, 50)), 3), "p99": round(float(np.percentile(a, 99)), 3)}


# ---------------------------------------------------------------------------
# GPU workloads
# ---------------------------------------------------------------------------
def bench_batch64(T, W, ctx, steps, warmup, batch_tiles):
    """Headline: 64 structures x 8 incidences, one batched mixed solve per
    step; plain FP64 one right-hand side at a time beside it."""
    probs = W.batch64(seed=4)
    incs = W.incidences()
    cores = [T.AProductCore(ctx, p.R, p.Nx, p.Ny, p.Nz, p.d, p.lam, p.nback) for p in probs]
    tuned = TUNE and batch_tiles == "tuned"
    if tuned:
        # each operator tuned on a scratch model of its 8 right-hand sides
        for c in cores:
            c.autotune(nrhs=len(incs))
    models, rhs_all = [], []
    for p, core in zip(probs, cores):
        rhs = [p.rhs(nK, nE)[:2] for nK, nE in incs]
        m = T.DeviceModel(core, len(incs))
        m.set(T.VEC_AL, np.concatenate([a for a, _ in rhs]))
        m.set(T.VEC_E, np.concatenate([e for _, e in rhs]))
        models.append(m)
        rhs_all.append(rhs)
    batch = T.Batch(models)
    nrhs = batch.size
    log(f"batch64: {len(probs)} structures, {nrhs} right-hand sides")
    for _ in range(warmup):
        batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
    ctx.sync()
    ts, ok, infos = [], True, None
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        infos = batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1)
        ts.append(time.perf_counter() - t1)
        ok = ok and converged(infos)
    ctx.sync()
    dt = time.perf_counter() - t0
    out = {"workload": "batch64: 64 structures, Nx,Ny in 16..64, Nz in 8..16, d=10nm, lambda=500nm, "
                       "eps 1.01..2.25+0.01i, 2 polarisations x 4 angles", "structures": len(probs), "rhs": nrhs,
           "dipoles_total": int(sum(p.N for p in probs)) * len(incs), "batch_tiles": "tuned" if tuned else "planner"}
    out["batched_mixed"] = {"solves_per_s": nrhs * steps / dt, "ms_per_step": dt / steps * 1e3,
                            "latency_ms": lat_stats(ts),
                            "iterations_mean": float(np.mean([i["iterations"] for i in infos])),
                            "outer_mean": float(np.mean([i["outer"] for i in infos])),
                            "true_residual_max": max(i["true_residual"] for i in infos), "all_converged": ok}
    # sample of right-hand sides for the field error: one per 8th structure
    sample = [(s, s % len(incs)) for s in range(0, len(probs), 8)]
    P_mixed = {(s, j): models[s].get(T.VEC_P, j) for s, j in sample}
    # the same batch with right Jacobi preconditioning (opt-in solver
    # option, not the reference's iteration): same system, same 1e-8 stop
    batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1, precond=T.PRECOND_JACOBI)
    ctx.sync()
    tsj, okj = [], True
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        jinfo = batch.solve(mode=T.MODE_MIXED, tol=TOL, x0_zero=1, true_residual=1, precond=T.PRECOND_JACOBI)
        tsj.append(time.perf_counter() - t1)
        okj = okj and converged(jinfo)
    ctx.sync()
    dtj = time.perf_counter() - t0
    out["batched_mixed_jacobi"] = {"solves_per_s": nrhs * steps / dtj, "ms_per_step": dtj / steps * 1e3,
                                   "latency_ms": lat_stats(tsj),
                                   "iterations_mean": float(np.mean([i["iterations"] for i in jinfo])),
                                   "outer_mean": float(np.mean([i["outer"] for i in jinfo])),
                                   "true_residual_max": max(i["true_residual"] for i in jinfo),
                                   "all_converged": okj}
    P_jac = {(s, j): models[s].get(T.VEC_P, j) for s, j in sample}
    # batched FP64
    batch.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1)
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        fin = batch.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1)
    dt64 = time.perf_counter() - t0
    out["batched_fp64"] = {"solves_per_s": nrhs * reps / dt64, "ms_per_step": dt64 / reps * 1e3,
                           "iterations_mean": float(np.mean([i["iterations"] for i in fin])),
                           "true_residual_max": max(i["true_residual"] for i in fin), "all_converged": converged(fin)}
    # operator passes over the batch (complex64: the inner solver's; complex128)
    cr = [(core, m.nrhs) for core, m in zip(cores, models)]
    for m in models:
        m.set(T.VEC_P, m.get(T.VEC_E))
    seq64 = batch.time_passes(T.C64, reps=10)
    seq128 = batch.time_passes(T.C128, reps=10)
    alone = batch.time_pass(2, T.C64, reps=20)
    b64, b128 = batch_bytes(cr, 8), batch_bytes(cr, 16)
    out["aproduct_passes"] = {"c64": pass_report(seq64, b64), "c128": pass_report(seq128, b128)}
    groups = len({c.grid[0] for c in cores})
    roof = roofline("k_px_xgreen", f"k_pxr<float, Mx, nt> ({groups} FFT-length groups, one launch each)", b64[2],
                    seq64[2], alone, "batch64_c64:k_px_xgreen",
                    "complex64 x pass over all 512 right-hand sides (R = 8 per structure, Green octant once per "
                    "structure); duration = all launches of the pass")
    batch.close()
    for m in models:
        m.close()
    # plain FP64: one right-hand side at a time, each operator with its
    # autotuned tiles (inputs resident; uploads between solves untimed)
    # (the Jacobi-preconditioned FP64 solve of each right-hand side timed
    # beside it)
    lat, pinfo, P_plain = [], [], {}
    latj, pinfoj = [], []
    for s, (core, rhs) in enumerate(zip(cores, rhs_all)):
        if TUNE and not tuned:
            core.autotune()
        one = T.DeviceModel(core, 1)
        for j, (a, e) in enumerate(rhs):
            one.set(T.VEC_AL, a)
            one.set(T.VEC_E, e)
            t0 = time.perf_counter()
            pinfoj.append(one.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1,
                                    precond=T.PRECOND_JACOBI)[0])
            latj.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            pinfo.append(one.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1)[0])
            lat.append(time.perf_counter() - t0)
            if (s, j) in P_mixed:
                P_plain[(s, j)] = one.get(T.VEC_P)[0]
        one.close()
    for key, ls, inf in (("plain_fp64", lat, pinfo), ("plain_fp64_jacobi", latj, pinfoj)):
        out[key] = {"solves_per_s": len(ls) / sum(ls), "latency_ms": lat_stats(ls),
                    "iterations_mean": float(np.mean([i["iterations"] for i in inf])),
                    "true_residual_max": max(i["true_residual"] for i in inf),
                    "all_converged": converged(inf)}
    out["batched_mixed_vs_plain_fp64"] = round(out["batched_mixed"]["solves_per_s"]
                                               / out["plain_fp64"]["solves_per_s"], 3)
    out["batched_fp64_vs_plain_fp64"] = round(out["batched_fp64"]["solves_per_s"]
                                              / out["plain_fp64"]["solves_per_s"], 3)
    out["batched_mixed_jacobi_vs_plain_fp64"] = round(out["batched_mixed_jacobi"]["solves_per_s"]
                                                      / out["plain_fp64"]["solves_per_s"], 3)
    out["batched_mixed_jacobi_vs_plain_fp64_jacobi"] = round(out["batched_mixed_jacobi"]["solves_per_s"]
                                                             / out["plain_fp64_jacobi"]["solves_per_s"], 3)
    cpu_cases = [dict(name=f"s{s}/rhs{j}", R=probs[s].R, dims=(probs[s].Nx, probs[s].Ny, probs[s].Nz),
                      d=probs[s].d, K=probs[s].K, grid=cores[s].grid, al=rhs_all[s][j][0], E=rhs_all[s][j][1],
                      gpu={"batched_mixed": P_mixed[(s, j)], "plain_fp64": P_plain[(s, j)],
                           "batched_mixed_jacobi": P_jac[(s, j)]}) for s, j in sample]
    for c in cores:
        c.close()
    return out, roof, cpu_cases


def bench_mixed128(T, W, ctx, reps):
    """128 x 128 x 32, 50 % occupancy binary design: mixed-precision solve
    (complex64 inner, FP64 refinement) against the plain FP64 solver; the
    operator's x pass in both precisions."""
    prob = W.mixed128(seed=3)
    al, E, _ = prob.rhs(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
    core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam, prob.nback)
    if TUNE:
        core.autotune()
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    out = {"workload": "mixed128: 128x128x32, d=8nm, lambda=600nm, 50% occupancy, eps=(2+0.01i)^2",
           "grid": list(core.grid), "dipoles": prob.N,
           "xpass": {"c128": core.xpass(T.C128), "c64": core.xpass(T.C64)},
           "ypass": {"c128": core.ypass(T.C128), "c64": core.ypass(T.C64)}}
    P = {}
    for mode, name, pc in ((T.MODE_FP64, "fp64", T.PRECOND_NONE), (T.MODE_MIXED, "mixed", T.PRECOND_NONE),
                           (T.MODE_FP64, "fp64_jacobi", T.PRECOND_JACOBI),
                           (T.MODE_MIXED, "mixed_jacobi", T.PRECOND_JACOBI)):
        m.solve(mode=mode, tol=TOL, x0_zero=1, true_residual=1, precond=pc)
        lat, infos = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            infos.append(m.solve(mode=mode, tol=TOL, x0_zero=1, true_residual=1, precond=pc)[0])
            lat.append(time.perf_counter() - t0)
        P[name] = m.get(T.VEC_P)[0]
        out[name] = {"solves_per_s": reps / sum(lat), "latency_ms": lat_stats(lat),
                     "iterations": infos[-1]["iterations"], "outer": infos[-1]["outer"],
                     "true_residual_max": max(i["true_residual"] for i in infos), "all_converged": converged(infos)}
    out["mixed_vs_plain_fp64"] = round(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    out["mixed_jacobi_vs_plain_fp64"] = round(out["mixed_jacobi"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    m.set(T.VEC_P, E)
    seq = {p: m.time_passes(p, reps=20) for p in (T.C128, T.C64)}
    b = {T.C128: pass_bytes(core, 16), T.C64: pass_bytes(core, 8)}
    out["aproduct_passes"] = {"c128": pass_report(seq[T.C128], b[T.C128]), "c64": pass_report(seq[T.C64], b[T.C64])}
    roofs = {}
    for p, key in ((T.C128, "c128"), (T.C64, "c64")):
        alone = m.time_pass(2, p, reps=200)
        kern = "k_pxr" if core.xpass(p) == "register" else "k_px"
        roofs[key] = roofline("k_px_xgreen", f"{kern}<{'double' if p == T.C128 else 'float'}, {core.grid[0]}>",
                              b[p][2], seq[p][2], alone, f"mixed128_{key}:k_px_xgreen",
                              "x pass of one 256 x 256 x 64 grid")
    m.close()
    core.close()
    case = dict(name="mixed128", R=prob.R, dims=(prob.Nx, prob.Ny, prob.Nz), d=prob.d, K=prob.K, grid=core.grid,
                al=al, E=E, gpu={k: P[k] for k in ("fp64", "mixed", "mixed_jacobi")})
    return out, roofs, case


def run_step(T, F, A, targets, diel, mode, lat, pc=0):
    """One topology step: forward solve, adjoint source, adjoint solve."""
    t0 = time.perf_counter()
    fi = F.solve(mode=mode, tol=TOL, x0_zero=1, true_residual=1, precond=pc)[0]
    t1 = time.perf_counter()
    idx = (3 * targets[:, None] + np.arange(3)[None, :]).reshape(-1)
    P = F.get_entries(T.VEC_P, 0, idx)
    al_t = F.al_host[idx]
    g = np.conj(al_t * P * 3.0 / (2.0 + diel[idx]))
    A.set_entries(T.VEC_E, 0, idx, g)
    ai = A.solve(mode=mode, tol=TOL, x0_zero=1, true_residual=1, precond=pc)[0]
    t2 = time.perf_counter()
    lat.append(t1 - t0)
    lat.append(t2 - t1)
    return fi, ai


def bench_topo80(T, W, ctx, steps, warmup):
    """Topology-optimisation step: forward solve (z-incident, x-polarised)
    and the adjoint solve whose source is conj(E_int) at 100 seeded target
    dipoles -- the two solves of EvoDDAModel::EvoOptimization
    (EvoDDAModel.cpp:435-437, 596-608) -- in both solver modes."""
    prob, targets = W.topo80(seed=2)
    al, E, diel = prob.rhs(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
    core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam, prob.nback)
    if TUNE:
        core.autotune()
    F = T.DeviceModel(core, 1)
    A = T.DeviceModel(core, 1)
    for m in (F, A):
        m.set(T.VEC_AL, al)
        m.al_host = al
    F.set(T.VEC_E, E)
    out = {"workload": "topo80: 80x80x16 slab, d=10nm, lambda=542nm, forward + adjoint "
                       "(100 targets) per step", "grid": list(core.grid), "dipoles": prob.N,
           "xpass_c128": core.xpass(T.C128), "ypass_c128": core.ypass(T.C128)}
    P = {}
    for mode, name, pc in ((T.MODE_FP64, "fp64", T.PRECOND_NONE), (T.MODE_MIXED, "mixed", T.PRECOND_NONE),
                           (T.MODE_MIXED, "mixed_jacobi", T.PRECOND_JACOBI)):
        lat = []
        for _ in range(warmup):
            run_step(T, F, A, targets, diel, mode, lat, pc)
        lat, infos = [], []
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            infos.append(run_step(T, F, A, targets, diel, mode, lat, pc))
        ctx.sync()
        dt = time.perf_counter() - t0
        out[name] = {"solves_per_s": 2 * steps / dt, "ms_per_step": dt / steps * 1e3, "latency_ms": lat_stats(lat),
                     "iterations_fwd": infos[-1][0]["iterations"], "iterations_adj": infos[-1][1]["iterations"],
                     "true_residual_max": max(i["true_residual"] for pair in infos for i in pair),
                     "all_converged": all(converged(pair) for pair in infos)}
        P[name] = (F.get(T.VEC_P)[0], A.get(T.VEC_E)[0], A.get(T.VEC_P)[0])
    out["mixed_vs_plain_fp64"] = round(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    out["mixed_jacobi_vs_plain_fp64"] = round(out["mixed_jacobi"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    F.set(T.VEC_P, E)
    out["aproduct_passes"] = {"c128": pass_report(F.time_passes(T.C128, reps=50), pass_bytes(core, 16)),
                              "c64": pass_report(F.time_passes(T.C64, reps=50), pass_bytes(core, 8))}
    F.close()
    A.close()
    core.close()
    base = dict(R=prob.R, dims=(prob.Nx, prob.Ny, prob.Nz), d=prob.d, K=prob.K, grid=core.grid, al=al)
    cases = [dict(base, name="topo80 forward", E=E, gpu={k: v[0] for k, v in P.items()}),
             dict(base, name="topo80 adjoint", E=P["fp64"][1], gpu={"fp64": P["fp64"][2]})]
    return out, cases


def bench_single16(T, W, ctx, reps):
    """16 x 16 x 16, eps = 2.25 everywhere, z-incident x-polarised plane
    wave: the reference algorithm in complex128, one solve at a time."""
    prob = W.single16()
    al, E, _ = prob.rhs(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))
    core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam, prob.nback)
    if TUNE:
        core.autotune()
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    m.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1)
    lat, infos = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        infos.append(m.solve(mode=T.MODE_FP64, tol=TOL, x0_zero=1, true_residual=1)[0])
        lat.append(time.perf_counter() - t0)
    P = m.get(T.VEC_P)[0]
    out = {"workload": "single16: 16x16x16, d=10nm, lambda=500nm, eps=2.25+0i, grid 32x32x32, complex128",
           "fp64": {"solves_per_s": reps / sum(lat), "latency_ms": lat_stats(lat),
                    "iterations": infos[-1]["iterations"],
                    "true_residual_max": max(i["true_residual"] for i in infos), "all_converged": converged(infos)}}
    m.close()
    core.close()
    case = dict(name="single16", R=prob.R, dims=(16, 16, 16), d=prob.d, K=prob.K, grid=core.grid, al=al, E=E,
                gpu={"fp64": P})
    return out, case


def bench_design(T, W, ctx, reps):
    """Design step of the topology workload on the device (tdda_design_step,
    the reference's CoreStructure::UpdateStr + DDAModel::UpdateAlpha): add a
    step to the 80 x 80 column parameters, cone filter r = 2, SmoothDensity
    (beta 8, eta 0.5), 1/alpha of 2 right-hand sides (complex128 and its
    complex64 copy).  One call moves the step in and runs three kernels."""
    prob, _ = W.topo80()
    core = T.AProductCore(ctx, prob.R, prob.Nx, prob.Ny, prob.Nz, prob.d, prob.lam, prob.nback)
    incs = [(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])), (np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0]))]
    m = T.DeviceModel(core, len(incs))
    ncol = prob.Nx * prob.Ny
    para0 = prob.t[::3][::prob.Nz].copy()
    dz = T.Design(m, para0, 2.0, prob.material, incs)
    rng = np.random.default_rng(7)
    steps = [0.01 * rng.standard_normal(ncol) for _ in range(max(reps, 1) * 20)]
    dz.step(steps[0], beta=8.0, eta=0.5)
    t0 = time.perf_counter()
    for st in steps:
        dz.step(st, beta=8.0, eta=0.5)
    dt = (time.perf_counter() - t0) / len(steps)
    n3 = 3 * prob.N
    # step in, origin read + write, filtered + para written, per column;
    # 1/alpha written per entry and right-hand side (16 B + the 8 B copy)
    algo = ncol * 8 * 5 + len(incs) * n3 * 24
    out = {"workload": "design: topo80 columns 80x80 (extruded over 16), cone filter r=2, beta=8, eta=0.5, 2 RHS",
           "columns": ncol, "rhs": len(incs), "step_us": round(dt * 1e6, 1),
           "algorithmic_bytes": int(algo), "effective_gbs": round(algo / dt / 1e9, 1)}
    # objective value and its adjoint source on the device (PointE at a point
    # above the slab; the source goes to the second right-hand side)
    m.set(T.VEC_P, np.tile(prob.rhs(*incs[0])[1], len(incs)))
    r0 = np.array([400.0, 400.0, 300.0])
    T.objective_point_e(m, 0, r0, incs[0][0], incs[0][1], 1.0, source=m, source_rhs=1)
    t0 = time.perf_counter()
    for _ in range(20):
        T.objective_point_e(m, 0, r0, incs[0][0], incs[0][1], 1.0, source=m, source_rhs=1)
    out["point_e_objective_and_source_us"] = round((time.perf_counter() - t0) / 20 * 1e6, 1)
    dz.close()
    m.close()
    core.close()
    return out


def bench_periodic(T, W, ctx, reps):
    """Periodic operator set-up on the device (tdda_op_create_periodic): the
    image-lattice tensor of an 80 x 80 x 16 unit cell (the planner's grid)
    with 21 x 21 images, then its spectrum -- the reference sums the images
    on one host thread (AProductCore.cpp:301-536)."""
    pm = pn = 10
    Nx, Ny, Nz, d, lam = 80, 80, 16, 10.0, 542.0
    T.AProductCore(ctx, None, 8, 8, 4, d, lam, periodic=(1, 1, 80.0, 80.0)).close()  # warm
    ts = []
    for _ in range(max(reps, 1)):
        t0 = time.perf_counter()
        core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, periodic=(pm, pn, Nx * d, Ny * d))
        ts.append(time.perf_counter() - t0)
        grid = core.grid
        core.close()
    evals = (2 * Nx - 1) * (2 * Ny - 1) * (2 * Nz - 1) * (2 * pm + 1) * (2 * pn + 1)
    return {"workload": "periodic: 80x80x16 cell, 21x21 images, Lm = Ln = 800 nm",
            "grid": list(grid), "images": (2 * pm + 1) * (2 * pn + 1), "setup_ms": round(float(np.median(ts)) * 1e3, 1),
            "tensor_evals_per_s": round(evals / float(np.median(ts)), 0)}


def bench_serve(T, seconds, rate=200.0, seed=5, precond=0):
    """Request stream (workloads.serve_stream): Poisson arrivals at
    `rate`/s for `seconds`, 8k-2M dipoles, 1-4 right-hand sides, half
    adjoint; mixed precision to a true residual of 1e-8 (with right Jacobi
    preconditioning when `precond` is PRECOND_JACOBI).  Latency = result
    ready - scheduled arrival (monotonic clock); service = result ready -
    admission."""
    import threading
    import workloads as W
    reqs, pool = W.serve_stream(seconds, rate, seed)
    mat = T.material_rule(np.array([1.0, 2.25 + 0.01j]))
    srv = T.Server(0, precond=precond)
    done = []
    tickets = []
    lock = threading.Condition()

    def collect():
        k = 0
        while True:
            with lock:
                while k >= len(tickets):
                    lock.wait()
                tk, at = tickets[k]
            k += 1
            if tk is None:
                return
            while srv.wait(tk, 10.0) == 0:
                # progress line for long runs
                log(f"serve: {len(done)} of {len(reqs)} requests done")
            try:
                _, infos, tm = srv.result(tk, want_P=False)
                done.append((at, tm, infos))
            except T.TddaError as e:
                done.append((at, None, str(e)))

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.monotonic()
    lag = []
    for r in reqs:
        now = time.monotonic()
        if t0 + r["at"] > now:
            time.sleep(t0 + r["at"] - now)
        lag.append(time.monotonic() - (t0 + r["at"]))
        tk = srv.submit(r["Nx"], r["Nx"], r["Nz"], 10.0, r["lam"], pool[r["off"]:r["off"] + 3 * r["N"]], mat,
                        r["incs"], sources=r["src"])
        with lock:
            tickets.append((tk, t0 + r["at"]))
            lock.notify()
    with lock:
        tickets.append((None, None))
        lock.notify()
    th.join()
    t_end = max(tm["t_done"] for _, tm, _ in done if tm)
    st = srv.stats()
    srv.close()
    ok = [(at, tm, infos) for at, tm, infos in done if tm]
    lat = np.array([tm["t_done"] - at for at, tm, _ in ok])
    svc = np.array([tm["t_done"] - tm["t_admit"] for _, tm, _ in ok])
    nrhs = sum(len(infos) for _, _, infos in ok)
    tr = [i["true_residual"] for _, _, infos in ok for i in infos]
    return {"workload": f"serve: Poisson {rate:g}/s for {seconds:g} s, dipoles log-U[8k, 2M], Nz in {{8,16,32}}, "
                        "1-4 RHS, half adjoint, mixed precision to 1e-8"
                        + (", right Jacobi preconditioning" if precond else ""),
            "requests": len(reqs), "completed": len(ok), "failed": len(done) - len(ok),
            "errors": dict(collections.Counter(err for _, tm, err in done if not tm).most_common(5)),
            "rhs_solved": nrhs, "solves_per_s": nrhs / (t_end - t0),
            "dipole_solves_per_s": sum(r["N"] * len(r["incs"]) for r in reqs) / (t_end - t0) if not
            (len(done) - len(ok)) else None,
            "requests_per_s": len(ok) / (t_end - t0), "wall_s": t_end - t0,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
            "service_ms": {"p50": float(np.percentile(svc, 50) * 1e3), "p99": float(np.percentile(svc, 99) * 1e3)},
            "submit_lag_ms_max": float(max(lag) * 1e3),
            "true_residual_max": max(tr), "all_converged": all(i["status"] == T.CONVERGED
                                                              for _, _, infos in ok for i in infos),
            "dipoles_mean": float(np.mean([r["N"] for r in reqs])),
            "peak_hbm_gb": st["peak_bytes"] / 1e9, "budget_gb": st["budget_bytes"] / 1e9, "ticks": st["ticks"],
            "worker_s": {"setup": st["setup_s"], "ticks": st["tick_s"], "release": st["release_s"]}}


# ---------------------------------------------------------------------------
# CPU legs: the oracle's restatement of the reference path on the host
# ---------------------------------------------------------------------------
def cpu_model():
    try:
        outp = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in outp.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def apply_ms(op, E, al, n=3):
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        op.apply(E, al)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3


def next_smooth(n):
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def cpu_case(O, c, allcores, one_core_full=False):
    """One right-hand side on the CPU: the oracle's BiCGSTAB to 1e-8 from
    zero on all host threads (timed), continued from there to 1e-11 (the
    reference solution for the field errors), the Aproduct on all threads
    and on one; `one_core_full` also times the 1e-8 solve on one thread."""
    Nx, Ny, Nz = c["dims"]
    # the port's own best grid: the smallest 2,3,5-smooth sizes >= 2N - 1
    # (the device may pick a longer x length whose transform is cheaper
    # there; the convolution is the same)
    grid = tuple(next_smooth(2 * n - 1) for n in (Nx, Ny, Nz))
    op = O.OrcOperator(c["R"], Nx, Ny, Nz, c["d"], c["K"], grid=grid)
    al, E = c["al"], c["E"]
    O.lib().orc_set_num_threads(allcores)
    t0 = time.perf_counter()
    P, st, it, _ = op.bicgstab(al, E, max_it=20000, tol=TOL)
    t_all = time.perf_counter() - t0
    Pref, st2, it2, _ = op.bicgstab(al, E, P0=P, max_it=20000, tol=REF_TOL)
    ref_res = float(np.linalg.norm(E - op.apply(Pref, al)) / np.linalg.norm(E))
    a_all = apply_ms(op, E, al)
    O.lib().orc_set_num_threads(1)
    a_one = apply_ms(op, E, al, n=1 if Nx * Ny * Nz > 200_000 else 3)
    t_one, one_kind = None, "projected: (2 iterations + 1) x Aproduct time on one thread"
    if one_core_full:
        t0 = time.perf_counter()
        op.bicgstab(al, E, max_it=20000, tol=TOL)
        t_one, one_kind = time.perf_counter() - t0, "measured"
    O.lib().orc_set_num_threads(allcores)
    if t_one is None:
        t_one = (2 * it + 1) * a_one * 1e-3
    errs = {k: O.rel_l2(v, Pref) for k, v in c["gpu"].items()}
    op.close()
    return {"name": c["name"], "grid": list(grid), "iterations": it, "status": st, "s_all": t_all,
            "aproduct_ms_all": a_all, "aproduct_ms_one": a_one, "s_one": t_one, "one_core": one_kind,
            "ref_iterations": it + it2, "ref_true_residual": ref_res, "field_error": errs}


def cpu_legs(cases, allcores):
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    res = {}
    for key, cs in cases.items():
        res[key] = [cpu_case(O, c, allcores, one_core_full=(key == "single16")) for c in cs]
        log(f"cpu: {key} done")
    return res


def summarise_cpu(legs, allcores):
    """Per-workload CPU rates: all host threads (measured) and one thread
    (measured for single16, projected otherwise)."""
    out = {}
    for key, rs in legs.items():
        s_all = sum(r["s_all"] for r in rs)
        s_one = sum(r["s_one"] for r in rs)
        out[key] = {"solves_per_s_all_cores": len(rs) / s_all, "solves_per_s_one_core": len(rs) / s_one,
                    "one_core": rs[0]["one_core"], "sample_solves": len(rs),
                    "iterations": [r["iterations"] for r in rs],
                    "aproduct_ms_all_cores": [round(r["aproduct_ms_all"], 2) for r in rs],
                    "aproduct_ms_one_core": [round(r["aproduct_ms_one"], 2) for r in rs],
                    "grid": rs[0]["grid"]}
    return out


def field_errors(legs):
    out = {}
    for key, rs in legs.items():
        e = collections.defaultdict(float)
        for r in rs:
            for k, v in r["field_error"].items():
                e[k] = max(e[k], v)
        out[key] = {"vs": "CPU oracle BiCGSTAB converged to 1e-11 (recurrence residual)",
                    "reference_true_residual_max": max(r["ref_true_residual"] for r in rs),
                    "max_rel_l2": dict(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workloads", default="single16,topo80,mixed128,design,periodic,serve,serve_jacobi",
                    help="workloads reported beside the headline (comma list, or none)")
    ap.add_argument("--extra-reps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="arrival window of the request stream")
    ap.add_argument("--no-tune", action="store_true", help="planner tiles instead of the autotuner's")
    ap.add_argument("--batch-tiles", default="planner", choices=["planner", "tuned"],
                    help="tiles of the batched solves (the single-RHS comparison always uses the autotuner's)")
    args = ap.parse_args()
    global TUNE
    TUNE = not args.no_tune
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): several GPUs would each take whole "
              "structures of the batch, one process per GPU, and that path is not built.",
              file=sys.stderr)
        sys.exit(2)

    import topodda as T
    import workloads as W

    ctx = T.Context(0)
    head, roof, cases4 = bench_batch64(T, W, ctx, args.steps, args.warmup, args.batch_tiles)
    log(f"batch64: {head['batched_mixed']['solves_per_s']:.1f} solves/s batched mixed, "
        f"{head['plain_fp64']['solves_per_s']:.1f} plain FP64")
    cpu_cases = {"batch64": cases4}
    extras, roofs = {}, {"batch64_c64": roof}
    for w in [x for x in args.workloads.split(",") if x and x != "none"]:
        log(f"workload {w}")
        if w == "serve":
            extras[w] = bench_serve(T, args.serve_seconds)
        elif w == "serve_jacobi":
            extras[w] = bench_serve(T, args.serve_seconds, precond=T.PRECOND_JACOBI)
        elif w == "mixed128":
            extras[w], r3, c3 = bench_mixed128(T, W, ctx, args.extra_reps)
            roofs["mixed128_c128"], roofs["mixed128_c64"] = r3["c128"], r3["c64"]
            cpu_cases[w] = [c3]
        elif w == "topo80":
            extras[w], c2 = bench_topo80(T, W, ctx, 10, 2)
            cpu_cases[w] = c2
        elif w == "single16":
            extras[w], c1 = bench_single16(T, W, ctx, 10)
            cpu_cases[w] = [c1]
        else:
            extras[w] = {"design": bench_design, "periodic": bench_periodic}[w](T, W, ctx, args.extra_reps)
    ctx.close()

    cpu, errs = None, None
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        allcores = O.lib().orc_num_threads()
        legs = cpu_legs(cpu_cases, allcores)
        per = summarise_cpu(legs, allcores)
        errs = field_errors(legs)
        b = per["batch64"]
        cpu = {"value": round(b["solves_per_s_all_cores"], 4), "unit": "solves/s", "cores": allcores, "kind": "port",
               "sample": f"batch64: {b['sample_solves']} of the 512 right-hand sides (one per 8th structure) solved "
                         "one at a time to 1e-8 by the oracle's C restatement (OpenMP FFT convolution, the "
                         "reference's BiCGSTAB) on all host threads; rate over that sample",
               "cpu": cpu_model(), "nproc": os.cpu_count(), "workloads": per}

    bm = head["batched_mixed"]
    line = {
        "metric": METRIC,
        "value": round(bm["solves_per_s"], 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bm["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128",
        "data": "synthetic",
        "config": {"workload": "batch64: 64 structures (Nx,Ny in 16..64, Nz in 8..16) x 8 incidences = 512 RHS in "
                               "one batch; complex64 inner BiCGSTAB + FP64 refinement to true residual 1e-8",
                   "structures": head["structures"], "rhs": head["rhs"], "dipoles_total": head["dipoles_total"],
                   "batch_tiles": head["batch_tiles"], "autotuned_single_rhs": TUNE},
        "latency_ms": bm["latency_ms"],
        "plain_fp64": head["plain_fp64"],
        "batched_fp64": head["batched_fp64"],
        "batched_mixed_vs_plain_fp64": head["batched_mixed_vs_plain_fp64"],
        "batched_mixed": bm,
        "jacobi": {"note": "opt-in right Jacobi preconditioning (BiCGSTAB on A diag(al)^-1, not the reference's "
                           "iteration): same systems, same 1e-8 true-residual stop",
                   "batched_mixed": head["batched_mixed_jacobi"], "plain_fp64": head["plain_fp64_jacobi"],
                   "batched_mixed_vs_plain_fp64": head["batched_mixed_jacobi_vs_plain_fp64"],
                   "batched_mixed_vs_plain_fp64_jacobi": head["batched_mixed_jacobi_vs_plain_fp64_jacobi"]},
        "aproduct_passes": head["aproduct_passes"],
        "roofline": roof,
        "roofline_by_workload": roofs,
        "field_error": errs,
        "cpu_baseline": cpu,
        "workloads": extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
