"""Writes the golden fixtures under tests/golden/ (run from the repo root).

The reference (Topo-DDA-forWin64) ships no tests, golden vectors or expected
outputs for this path, so the fixtures come from:

* alpha_ref.npz -- the reference's own Get_Alpha (tools.cpp:617-630),
  nEPerpinXZ (tools.cpp:686-709) and build_a_bulk (tools.cpp:497-510),
  compiled from its sources into oracle/_ref/libref_tools.so
  (oracle/Makefile target ``ref``).
* green_ld.npz -- the interaction tensor at lattice offsets, evaluated in
  long double with the textbook dyadic Green function -G (the algebraic
  content of AProductCore.cpp:588-623).
* aproduct_tiny.npz -- off-diagonal products y_i = sum_{j != i} A(r_i - r_j) b_j
  by long-double direct summation on small lattices with random occupancy
  (what AProductCore::Aproduct, AProductCore.cpp:697-758, computes).
* solve_tiny.npz -- dense numpy.linalg.solve of the assembled 3N x 3N system
  A P = E of DDAModel (DDAModel.cpp:12-69, 577-586) for small structures.
* design_ref.npz -- the reference's own filtered design update of an extruded
  box (CoreStructure::UpdateStr, CoreStructure.cpp:31-104, through SpacePara's
  FreeWeight and FilterOption::SmoothDensity) and its beta schedules
  (tools.cpp:776-798), compiled from its sources into oracle/_ref.

``python tests/golden/make_golden.py`` writes every fixture;
``python tests/golden/make_golden.py design`` only design_ref.npz.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def alpha_fixture():
    ref = O.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_tools.so missing: run make -C oracle ref")
    rng = np.random.default_rng(11)
    rows = []
    thetas = np.deg2rad([0.0, 10.0, 20.0, 30.0, 45.0, 60.0, 85.0])
    for lam in (450.0, 500.0, 542.0, 600.0, 700.0):
        for d in (2.0, 8.0, 10.0):
            for diel in (1.01 + 0j, 2.25 + 0j, 2.25 + 0.01j, 3.9999 + 0.04j, 1.6 + 0.005j,
                         complex(rng.uniform(1.01, 4), rng.uniform(0, 0.1))):
                th = rng.choice(thetas)
                nK = np.array([np.sin(th), 0.0, np.cos(th)])
                nE = np.array([np.cos(th), 0.0, -np.sin(th)]) if rng.random() < 0.5 else np.array([0.0, 1.0, 0.0])
                rows.append((lam, d, diel, nE, nK))
    lam = np.array([r[0] for r in rows])
    d = np.array([r[1] for r in rows])
    diel = np.array([r[2] for r in rows])
    nE = np.array([r[3] for r in rows])
    nK = np.array([r[4] for r in rows])
    K = 2 * np.pi / lam
    alpha = np.zeros(len(rows), np.complex128)
    out = np.zeros(2)
    dp = O._dp
    for i in range(len(rows)):
        ref.ref_get_alpha(lam[i], K[i], d[i], diel[i].real, diel[i].imag,
                          O._ptr(np.ascontiguousarray(nE[i])), O._ptr(np.ascontiguousarray(nK[i])), O._ptr(out))
        alpha[i] = out[0] + 1j * out[1]
    nperp = np.zeros((len(thetas), 3))
    for i, th in enumerate(thetas):
        v = np.zeros(3)
        ref.ref_nE_perp_in_xz(th, 0.0, O._ptr(v))
        nperp[i] = v
    bulk = np.zeros(3 * 2 * 3 * 4, np.int32)
    ref.ref_build_a_bulk(2, 3, 4, O._ptr(bulk, O._ip))
    del dp
    np.savez(os.path.join(OUT, "alpha_ref.npz"), lam=lam, d=d, K=K, diel=diel, nE=nE, nK=nK, alpha=alpha,
             thetas=thetas, nperp=nperp, bulk234=bulk)


def green_fixture():
    K = 2 * np.pi / 500.0
    offs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 2, 3], [-3, 2, -1], [5, -4, 2], [-7, -7, 7],
                     [15, 0, -15], [31, 1, 2], [2, 2, 0]], float) * 10.0
    A = O.green_textbook(K, offs[:, 0], offs[:, 1], offs[:, 2], dtype=np.clongdouble)
    np.savez(os.path.join(OUT, "green_ld.npz"), K=K, offs=offs, A=A.astype(np.complex128))


def aproduct_fixture():
    rng = np.random.default_rng(5)
    cases = {}
    for name, (Nx, Ny, Nz, frac, d, lam) in {
        "a": (3, 4, 5, 0.7, 10.0, 500.0),
        "b": (5, 3, 2, 1.0, 7.0, 450.0),
        "c": (6, 6, 4, 0.5, 8.0, 600.0),
    }.items():
        R = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
        if frac < 1:
            R = R[rng.random(len(R)) < frac]
        R = R.reshape(-1).astype(np.int32)
        N = len(R) // 3
        K = 2 * np.pi / lam
        b = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
        y = O.aproduct_direct_ld(R, d, K, b).astype(np.complex128)
        cases.update({f"{name}_dims": np.array([Nx, Ny, Nz]), f"{name}_R": R, f"{name}_d": d, f"{name}_K": K,
                      f"{name}_lam": lam, f"{name}_b": b, f"{name}_y": y})
    np.savez(os.path.join(OUT, "aproduct_tiny.npz"), **cases)


def solve_fixture():
    rng = np.random.default_rng(9)
    cases = {}
    for name, (Nx, Ny, Nz, d, lam, th_deg, pol) in {
        "a": (4, 3, 3, 10.0, 500.0, 0.0, "x"),
        "b": (3, 3, 5, 10.0, 542.0, 20.0, "p"),
        "c": (5, 4, 2, 8.0, 600.0, 10.0, "y"),
    }.items():
        R = O.build_a_bulk(Nx, Ny, Nz)
        N = len(R) // 3
        th = np.deg2rad(th_deg)
        nK = np.array([np.sin(th), 0.0, np.cos(th)])
        nE = {"x": np.array([1.0, 0, 0]), "y": np.array([0, 1.0, 0]), "p": O.ne_perp_in_xz(th)}[pol]
        K = O.wavenumber(lam)
        mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
        t = rng.random(N).repeat(3)
        diel, al = O.diel_al(t, mat, lam, K, d, nE, nK)
        E = O.einc(R, K, d, 1.0, nK, nE)
        A = O.assemble_dense(R, d, K, al)
        P = np.linalg.solve(A, E)
        cases.update({f"{name}_dims": np.array([Nx, Ny, Nz]), f"{name}_d": d, f"{name}_lam": lam,
                      f"{name}_nK": nK, f"{name}_nE": nE, f"{name}_t": t, f"{name}_mat": mat,
                      f"{name}_al": al, f"{name}_E": E, f"{name}_P": P})
    np.savez(os.path.join(OUT, "solve_tiny.npz"), **cases)


# (Nx, Ny, Nz, rfilter, beta, eta, step scale): radii below one lattice step
# (self only), 2 and 3.3 (9 and 37 columns), beta 0 (identity projection)
# and steep projections, steps that cross the clamps
DESIGN_CASES = [(5, 4, 2, 2.0, 0.0, 0.5, 0.1), (7, 6, 1, 3.3, 20.0, 0.3, 0.2), (3, 9, 3, 0.9, 8.0, 0.5, 0.5),
                (6, 6, 2, 2.0, 50.0, 0.7, 0.8)]


def design_fixture():
    ref = O.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_tools.so missing: run make -C oracle ref")
    rng = np.random.default_rng(21)
    cases = {}
    for k, (Nx, Ny, Nz, r, beta, eta, sc) in enumerate(DESIGN_CASES):
        p0 = rng.random(Nx * Ny)
        st = sc * rng.standard_normal(Nx * Ny)
        diel = np.zeros(3 * Nx * Ny * Nz)
        f = np.zeros(Nx * Ny)
        p = np.zeros(Nx * Ny)
        rc = ref.ref_design_update(Nx, Ny, Nz, O._ptr(p0), O._ptr(st), r, beta, eta, O._ptr(diel), O._ptr(f),
                                   O._ptr(p))
        if rc != 0:
            raise SystemExit(f"reference design update failed: {rc}")
        cases.update({f"c{k}_case": np.array([Nx, Ny, Nz, r, beta, eta]), f"c{k}_p0": p0, f"c{k}_step": st,
                      f"c{k}_diel": diel, f"c{k}_filtered": f, f"c{k}_para": p})
    xs = np.array([0.0, 1.0, 2.5, 5.0, 6.0, 7.5, 8.0, 9.9, 10.0])
    cases["beta_x"] = xs
    cases["beta"] = np.array([[ref.ref_beta_update(kind, x, 10.0, 1.0, 200.0) for x in xs] for kind in range(3)])
    np.savez(os.path.join(OUT, "design_ref.npz"), **cases)


if __name__ == "__main__":
    if sys.argv[1:] == ["design"]:
        design_fixture()
        raise SystemExit(0)
    alpha_fixture()
    green_fixture()
    aproduct_fixture()
    solve_fixture()
    design_fixture()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))
