"""CPU parity of the host design classes (SpacePara, FilterOption,
CoreStructure::UpdateStr; the reference's SpacePara.cpp:1087-1384 with
ChangeFilter, CoreStructure.cpp:6-104, FilterOption.cpp) against the
reference's own classes compiled from its sources (oracle/_ref): the same
design case through both gives the same parameter map, the same cone-filter
weights in the same order (periodic wrap, 4-fold symmetry, fixed background
parameters, a filter radius change) and the same design values bit for bit
after several update steps."""
import os
import subprocess

import numpy as np
import pytest

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
    return out, rng


def run_case(tmp_path, name, Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, nfree, seed=0):
    lines, rng = case_text(Nx, Ny, Nz, bindz, structs, filt, sym, periodic, steps, seed)
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


needs_ref = pytest.mark.skipif(O.ref_lib() is None or not os.path.exists(EXE),
                               reason="reference classes (oracle/_ref) or dda_design_check not built")


@needs_ref
def test_extruded_box_filter_change(tmp_path):
    Nx, Ny, Nz = 7, 6, 3
    run_case(tmp_path, "box", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 30.0, 0.5, "exp", [(0, 1.5), (2, 2.3)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 1), (1, 1), (2, 1), (3, 1)], Nx * Ny)


@needs_ref
def test_periodic_wrap(tmp_path):
    Nx, Ny, Nz = 8, 7, 2
    run_case(tmp_path, "periodic", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 2.0, 20.0, 0.4, "linear", [(0, 2.2)]), ("None", 0.0, 0.0), (True, Nx, Ny),
             [(0, 0), (1, 0), (2, 0)], Nx * Ny, seed=1)


@needs_ref
def test_periodic_period_shorter_than_filter(tmp_path):
    # periods of 3 with a radius of 2: a site's images fall in one disc
    # several times, each image adding its own weight
    Nx, Ny, Nz = 3, 3, 2
    run_case(tmp_path, "short", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 8.0, 0.5, "piecewise", [(0, 2.0)]), ("None", 0.0, 0.0), (True, Nx, Ny),
             [(0, 0), (1, 0), (2, 0), (3, 0)], Nx * Ny, seed=2)


@needs_ref
def test_four_fold_symmetry(tmp_path):
    Nx, Ny, Nz = 8, 8, 2
    run_case(tmp_path, "sym", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (True, 1.0, 10.0, 0.5, "linear", [(0, 1.8)]), ("4fold", 3.5, 3.5), (False, 0, 0),
             [(0, 0), (1, 0)], 16, seed=3)


@needs_ref
def test_free_layer_over_fixed_substrate(tmp_path):
    # a free top layer filtered over the fixed substrate's parameters too
    Nx, Ny = 6, 5
    free = bulk(range(Nx), range(Ny), [2, 3])
    fixed = bulk(range(Nx), range(Ny), [0, 1])
    run_case(tmp_path, "substrate", Nx, Ny, 4, 2, [(1, free), (0, fixed)],
             (True, 1.0, 5.0, 0.5, "exp", [(0, 1.6)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 0), (1, 0), (2, 0)], Nx * Ny, seed=4)


@needs_ref
def test_unfiltered_steps_clamp(tmp_path):
    Nx, Ny, Nz = 5, 4, 2
    run_case(tmp_path, "nofilter", Nx, Ny, Nz, Nz, [(1, bulk(range(Nx), range(Ny), range(Nz)))],
             (False, 0.0, 0.0, 0.5, "linear", [(0, 1.0)]), ("None", 0.0, 0.0), (False, 0, 0),
             [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0)], Nx * Ny, seed=5)
