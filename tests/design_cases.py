"""Design cases shared by the host-class and device design tests: a case
text (the input of dda_design_check and of the reference shim
ref_spacepara_case) and the comparison of the two runs."""
import os
import subprocess

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "topo-dda-forwin64_amd", "dda_design_check")


def bulk(xs, ys, zs):
    return [(x, y, z) for x in xs for y in ys for z in zs]


def case_text(Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, seed):
    """structs: [(para, sites)]; filt: (on, bmin, bmax, ita, type, [(it, r)]);
    sym: (name, ax, ay); periodic: (on, Lx, Ly); steps: [(it, change)]."""
    rng = np.random.default_rng(seed)
    sites = [s for _, ss in structs for s in ss]
    out = [f"{Nx} {Ny} {Nz} {bindz}", str(len(structs))]
    for para, ss in structs:
        out.append(f"{para} {len(ss)} " + " ".join(f"{x} {y} {z}" for x, y, z in ss))
    vals = rng.random(len(sites))
    out.append(str(len(sites)))
    out.append(" ".join(f"{x} {y} {z}" for x, y, z in sites))
    out.append(" ".join(repr(float(v)) for v in vals))
    on, bmin, bmax, ita, kind, fl = filt
    out.append(f"{int(on)} {bmin!r} {bmax!r} {ita!r} {kind} {len(fl)} " + " ".join(f"{i} {r!r}" for i, r in fl))
    out.append(f"{sym[0]} {sym[1]!r} {sym[2]!r}")
    out.append(f"{int(periodic[0])} {periodic[1]} {periodic[2]}")
    return out, rng, vals


def run_case(tmp_path, name, Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, nfree, seed=0):
    lines, rng, _ = case_text(Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, seed)
    lines.append(f"{len(steps)} {max(it for it, _ in steps) + 1}")
    for it, change in steps:
        lines.append(f"{it} {change} " + " ".join(repr(float(v)) for v in 0.2 * rng.standard_normal(nfree)))
    case = tmp_path / f"{name}.txt"
    case.write_text("\n".join(lines) + "\n")
    ours, ref = tmp_path / f"{name}_ours.txt", tmp_path / f"{name}_ref.txt"
    p = subprocess.run([EXE, str(case), str(ours)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rc = O.ref_lib().ref_spacepara_case(str(case).encode(), str(ref).encode())
    assert rc == 0
    a, b = ours.read_text().split(), ref.read_text().split()
    head = [int(x) for x in a[:3]]
    assert head[2] == nfree, head
    assert a == b, f"{name}: first difference at token {next(i for i, (x, y) in enumerate(zip(a, b)) if x != y)}"
    return head




def ref_case(tmp_path, name, Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, nfree, seed=0):
    """The case through the reference's classes only; returns the parsed
    output (geometryPara, weights, Para / Para_origin / Para_filtered after
    the steps), the initial design values per site and the steps."""
    lines, rng, vals = case_text(Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, seed)
    lines.append(f"{len(steps)} {max(it for it, _ in steps) + 1}")
    st = []
    for it, change in steps:
        v = 0.2 * rng.standard_normal(nfree)
        st.append((it, v))
        lines.append(f"{it} {change} " + " ".join(repr(float(x)) for x in v))
    case = tmp_path / f"{name}.txt"
    case.write_text("\n".join(lines) + "\n")
    ref = tmp_path / f"{name}_ref.txt"
    assert O.ref_lib().ref_spacepara_case(str(case).encode(), str(ref).encode()) == 0
    tok = ref.read_text().split()
    k = 0

    def take(n, f):
        nonlocal k
        out = [f(x) for x in tok[k:k + n]]
        k += n
        return out
    N, npara, nf = take(3, int)
    gp = np.array(take(N, int))
    weights = []
    for _ in range(nf if filt[0] else 0):
        (nw,) = take(1, int)
        pairs = take(2 * nw, float)
        weights.append([(int(pairs[2 * i + 1]), pairs[2 * i]) for i in range(nw)])
    para = np.array(take(npara, float))
    origin = np.array(take(npara, float)) if filt[0] else None
    filtered = np.array(take(npara, float)) if filt[0] else None
    return dict(N=N, npara=npara, nfree=nf, gp=gp, weights=weights, para=para, origin=origin, filtered=filtered,
                site_vals=np.asarray(vals), steps=st, max_it=max(it for it, _ in steps) + 1)
