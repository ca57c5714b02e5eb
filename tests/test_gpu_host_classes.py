"""GPU test of the C++ host classes (AProductCore, DDAModel, EvoDDAModel over
the C ABI): the dda_host_check driver runs a forward solve, an adjoint solve
(EvoDDAModel.cpp:596-608 order) and Aproductwithalb; results are compared with
the oracle's restatement of the same steps."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "topo-dda-forwin64_amd", "dda_host_check")


def test_host_classes_forward_adjoint(tmp_path):
    Nx, Ny, Nz, d, lam = 10, 8, 6, 10.0, 542.0
    out = tmp_path / "host.bin"
    print("dda_host_check 10x8x6", flush=True)
    r = subprocess.run([EXE, str(out), str(Nx), str(Ny), str(Nz), str(d), str(lam), "5"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    N, it_f, it_a, Mx, My, Mz = (int(v) for v in np.frombuffer(raw[:24], np.int32))
    n3 = 3 * N
    h = 24
    vecs = np.frombuffer(raw[h:h + 6 * n3 * 16], np.complex128).reshape(6, n3)
    Pf, lam_adj, y, al, E, diel = vecs
    nt = int(np.frombuffer(raw[h + 6 * n3 * 16:h + 4 + 6 * n3 * 16], np.int32)[0])
    targets = np.frombuffer(raw[h + 4 + 6 * n3 * 16:], np.int32)[:nt]
    R = O.build_a_bulk(Nx, Ny, Nz)
    K = O.wavenumber(lam)
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=(Mx, My, Mz))
    assert O.rel_l2(y, op.apply(E, al)) < 1e-12
    # solutions converged to 1e-10 agree to ~1e-9; iteration counts of these
    # non-normal systems move by a few under rounding differences
    Pc, st, it, _ = op.bicgstab(al, E, max_it=1000, tol=1e-12)
    print("iterations fwd", it_f, "oracle (1e-12)", it, flush=True)
    assert st == 0 and O.rel_l2(Pf, Pc) < 1e-8
    Eint = O.local_field(al, Pc, diel)
    devp = np.zeros(n3, complex)
    for t in targets:
        devp[3 * t:3 * t + 3] = np.conj(Eint[3 * t:3 * t + 3])
    Lc, st, it, _ = op.bicgstab(al, devp, max_it=1000, tol=1e-12)
    print("iterations adj", it_a, "oracle (1e-12)", it, flush=True)
    assert st == 0 and O.rel_l2(lam_adj, Lc) < 1e-7
    # reciprocity of the symmetric operator: lambda^T E = devp^T P
    assert abs(lam_adj @ E - devp @ Pf) < 1e-7 * abs(devp @ Pf)
    # text files of the forward solve (CoreStructure.cpp:165-190,
    # DDAModel.cpp:618-697; their bytes are pinned against the reference on
    # the CPU): the structure's integers exactly and its design values, EResult
    # (= Einternal) and P to the 6 significant digits the files hold
    st = (tmp_path / "CoreStructure0.txt").read_text().split("\n")
    assert [int(v) for v in st[:4]] == [Nx, Ny, Nz, N] and st[-1] == ""
    assert np.array_equal(np.array([int(v) for v in st[4:4 + n3]]), R)
    # diel = m0 + t (m1 - m0) with the material rule's m0 = 1 + 0.01, m1 = 2.25 + 0.01i
    t = (diel.real - 1.01) / 1.24
    assert np.all(np.abs(np.array([float(v) for v in st[4 + n3:4 + 2 * n3]]) - t) <= 5.01e-6 * np.abs(t) + 1e-12)
    assert float(st[4 + 2 * n3]) == d and len(st) == 4 + 2 * n3 + 2
    lines = (tmp_path / "Model_results0it0.txt").read_text().splitlines()
    assert len(lines) == 2 * n3
    _close6(_parse(lines[:n3]), O.local_field(al, Pf, diel))
    _close6(_parse(lines[n3:]), Pf)
    pts = np.array([-2, 0, 0, Nx + 1, 1, 0, 0, -3, 1, 1, Ny + 2, Nz - 1, 0, 0, -2, Nx // 2, Ny // 2, Nz + 3])
    ER = O.field_at_points(R, d, K, Pf, pts, [0, 0, 1], [1, 0, 0], 1.0).reshape(-1)
    _close6(_parse((tmp_path / "Model_resultsit1.txt").read_text().splitlines()), ER)


_NUM = r"[+-]?(?:\d+(?:\.\d*)?(?:e[+-]?\d+)?|inf|nan)"


def _parse(lines):
    out = []
    for ln in lines:
        m = re.fullmatch(f"({_NUM})\\+?({_NUM})j", ln)
        assert m, ln
        out.append(complex(float(m.group(1)), float(m.group(2))))
    return np.array(out)


def _close6(got, want):
    # "%g": 6 significant digits per real number
    for a, b in ((got.real, want.real), (got.imag, want.imag)):
        assert np.all(np.abs(a - b) <= 5.01e-6 * np.abs(b) + 1e-300)


@pytest.mark.parametrize("method", ["LDR", "FCD"])
def test_host_classes_periodic_overload(tmp_path, method):
    # the periodic AProductCore overload (AProductCore.cpp:251-536) through the
    # same driver: Aproductwithalb against the oracle's image sum, and the
    # forward solution's residual under that operator
    Nx, Ny, Nz, d, lam = 6, 5, 3, 10.0, 542.0
    images = (2, 1, 70.0, 60.0)
    out = tmp_path / "host_periodic.bin"
    print("dda_host_check 6x5x3 periodic", method, flush=True)
    r = subprocess.run([EXE, str(out), str(Nx), str(Ny), str(Nz), str(d), str(lam), "7"] + [str(v) for v in images] +
                       [method], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    N = int(np.frombuffer(raw[:4], np.int32)[0])
    n3 = 3 * N
    Pf, lam_adj, y, al, E, diel = np.frombuffer(raw[24:24 + 6 * n3 * 16], np.complex128).reshape(6, n3)
    R = O.build_a_bulk(Nx, Ny, Nz)
    K = O.wavenumber(lam)
    fcd = method == "FCD"
    assert O.rel_l2(y, O.aproduct_periodic_direct(R, d, K, E, *images, fcd=fcd) + al * E) < 1e-12
    res = O.aproduct_periodic_direct(R, d, K, Pf, *images, fcd=fcd) + al * Pf
    assert O.rel_l2(res, E) < 1e-9


def test_host_classes_fcd(tmp_path):
    # AProductCore(..., "FCD") (AProductCore.cpp:12-249 with FCD_inter,
    # 625-695) through the driver: Aproductwithalb against the oracle's direct
    # sum of the FCD tensor, and the forward solution's residual
    Nx, Ny, Nz, d, lam = 7, 5, 4, 10.0, 542.0
    out = tmp_path / "host_fcd.bin"
    print("dda_host_check 7x5x4 FCD", flush=True)
    r = subprocess.run([EXE, str(out), str(Nx), str(Ny), str(Nz), str(d), str(lam), "9", "FCD"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    N = int(np.frombuffer(raw[:4], np.int32)[0])
    n3 = 3 * N
    Pf, lam_adj, y, al, E, diel = np.frombuffer(raw[24:24 + 6 * n3 * 16], np.complex128).reshape(6, n3)
    R = O.build_a_bulk(Nx, Ny, Nz)
    K = O.wavenumber(lam)
    assert O.rel_l2(y, O.aproduct_fcd_direct(R, d, K, E) + al * E) < 1e-12
    assert O.rel_l2(O.aproduct_fcd_direct(R, d, K, Pf) + al * Pf, E) < 1e-9
