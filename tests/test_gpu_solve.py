"""GPU parity of DDAModel::bicgstab (DDAModel.cpp:131-260): the FP64 device
solver against the oracle's restatement (same iteration count, solution
within 1e-10 relative L2) and against dense solves; the mixed-precision
solver against the 1e-8 true-residual bar."""
import os

import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def make(Nx, Ny, Nz, d, lam, seed=0, th_deg=0.0, pol="x", mat=(1.0, 2.25 + 0.01j), t=None):
    R = O.build_a_bulk(Nx, Ny, Nz)
    N = len(R) // 3
    th = np.deg2rad(th_deg)
    nK = np.array([np.sin(th), 0.0, np.cos(th)])
    nE = {"x": np.array([1.0, 0, 0]), "y": np.array([0, 1.0, 0]), "p": O.ne_perp_in_xz(th)}[pol]
    K = O.wavenumber(lam)
    m = O.material_rule(np.array(mat, complex))
    if t is None:
        t = np.random.default_rng(seed).random(N).repeat(3)
    diel, al = O.diel_al(t, m, lam, K, d, nE, nK)
    E = O.einc(R, K, d, 1.0, nK, nE)
    return R, N, K, al, E, diel


def test_fp64_solve_matches_dense_fixture(ctx):
    g = np.load(os.path.join(GOLD, "solve_tiny.npz"), allow_pickle=False)
    for name in ("a", "b", "c"):
        Nx, Ny, Nz = (int(v) for v in g[f"{name}_dims"])
        print(f"dense fixture {name}", flush=True)
        core = T.AProductCore(ctx, O.build_a_bulk(Nx, Ny, Nz), Nx, Ny, Nz, float(g[f"{name}_d"]),
                              float(g[f"{name}_lam"]))
        m = T.DeviceModel(core, 1)
        m.set(T.VEC_AL, g[f"{name}_al"])
        m.set(T.VEC_E, g[f"{name}_E"])
        m.set(T.VEC_P, np.zeros_like(g[f"{name}_E"]))
        info = m.solve(max_it=2000, tol=1e-13)[0]
        assert info["status"] == T.CONVERGED
        assert O.rel_l2(m.get(T.VEC_P)[0], g[f"{name}_P"]) < 1e-11


CASES = [(16, 16, 16, 10.0, 500.0, 0.0, "x", True),
         (12, 10, 6, 10.0, 542.0, 20.0, "p", False),
         (9, 11, 5, 8.0, 600.0, 10.0, "y", False)]


@pytest.mark.parametrize("case", CASES)
def test_fp64_iterates_match_oracle(ctx, case):
    """Same algorithm statement by statement: the first iterates agree to
    rounding.  (BiCGSTAB on these non-normal systems amplifies last-bit
    differences after ~10 iterations -- the oracle run on two grid sizes
    diverges from 1e-14 at iteration 10 to 1e-7 at iteration 20 -- so later
    iterates are compared through the converged solution instead.)"""
    Nx, Ny, Nz, d, lam, th, pol, uniform = case
    print(f"fp64 iterates {case}", flush=True)
    # the uniform 16^3 box is the single-solve workload: eps = 2.25 + 0i everywhere
    t = np.ones(3 * Nx * Ny * Nz) if uniform else None
    mat = (1.0, 2.25) if uniform else (1.0, 2.25 + 0.01j)
    R, N, K, al, E, _ = make(Nx, Ny, Nz, d, lam, th_deg=th, pol=pol, t=t, seed=Nx, mat=mat)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
    rng = np.random.default_rng(1)
    x0 = 0.01 * (rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N))  # warm start: A x0 path
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    for k in (1, 2, 4, 8):
        m.set(T.VEC_P, x0)
        info = m.solve(max_it=k, tol=0.0)[0]
        Pc, st, it, err = op.bicgstab(al, E, P0=x0, max_it=k, tol=0.0)
        assert info["status"] == T.MAXIT and info["iterations"] == k and st == 2
        assert O.rel_l2(m.get(T.VEC_P)[0], Pc) < 1e-11
        assert abs(info["error"] - err) <= 1e-9 * err


@pytest.mark.parametrize("case", CASES)
def test_fp64_converged_solution_matches_oracle(ctx, case):
    Nx, Ny, Nz, d, lam, th, pol, uniform = case
    print(f"fp64 solve {case}", flush=True)
    # the uniform 16^3 box is the single-solve workload: eps = 2.25 + 0i everywhere
    t = np.ones(3 * Nx * Ny * Nz) if uniform else None
    mat = (1.0, 2.25) if uniform else (1.0, 2.25 + 0.01j)
    R, N, K, al, E, _ = make(Nx, Ny, Nz, d, lam, th_deg=th, pol=pol, t=t, seed=Nx, mat=mat)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    m.set(T.VEC_P, np.zeros_like(E))
    info = m.solve(max_it=1000, tol=1e-8, true_residual=1)[0]
    _, st, it, _ = op.bicgstab(al, E, max_it=1000, tol=1e-8)
    Pref, _, _, _ = op.bicgstab(al, E, max_it=5000, tol=1e-13)
    assert st == 0 and info["status"] == T.CONVERGED
    assert abs(info["iterations"] - it) <= max(2, it // 10)
    assert info["error"] <= 1e-8 and info["true_residual"] <= 1.0001e-8
    assert O.rel_l2(m.get(T.VEC_P)[0], Pref) < 1e-6


def test_breakdown_and_maxit_follow_reference(ctx):
    R, N, K, al, E, _ = make(5, 4, 3, 10.0, 500.0)
    core = T.AProductCore(ctx, R, 5, 4, 3, 10.0, 500.0)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    # zero right-hand side: breakdown at the first loop test, ITERATION = 1
    m.set(T.VEC_E, np.zeros_like(E))
    m.set(T.VEC_P, np.zeros_like(E))
    info = m.solve(max_it=100, tol=1e-8)[0]
    assert info["status"] == T.BREAKDOWN and info["iterations"] == 1
    # iteration cap
    m.set(T.VEC_E, E)
    m.set(T.VEC_P, np.zeros_like(E))
    info = m.solve(max_it=1, tol=1e-30)[0]
    assert info["status"] == T.MAXIT and info["iterations"] == 1


def test_mixed_precision_reaches_true_residual(ctx):
    for (Nx, Ny, Nz, d, lam) in [(16, 16, 16, 10.0, 500.0), (24, 20, 8, 10.0, 542.0)]:
        print(f"mixed {Nx}x{Ny}x{Nz}", flush=True)
        R, N, K, al, E, _ = make(Nx, Ny, Nz, d, lam, seed=3)
        core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
        op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
        Pref, st, _, _ = op.bicgstab(al, E, max_it=5000, tol=1e-12)
        m = T.DeviceModel(core, 1)
        m.set(T.VEC_AL, al)
        m.set(T.VEC_E, E)
        info = m.solve(mode=T.MODE_MIXED, tol=1e-8, x0_zero=1)[0]
        P = m.get(T.VEC_P)[0]
        assert info["status"] == T.CONVERGED
        true_res = O.rel_l2(op.apply(P, al), E)
        assert true_res <= 1e-8
        assert abs(true_res - info["true_residual"]) < 1e-10
        assert O.rel_l2(P, Pref) < 1e-6


def test_adjoint_solve_uses_same_operator(ctx):
    # EvoDDAModel.cpp:596-608: change_E(devp), solve with the same A (A^T = A)
    R, N, K, al, E, diel = make(8, 8, 4, 10.0, 542.0, seed=5)
    core = T.AProductCore(ctx, R, 8, 8, 4, 10.0, 542.0)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    m.set(T.VEC_P, np.zeros_like(E))
    m.solve(tol=1e-12)
    P = m.get(T.VEC_P)[0]
    rng = np.random.default_rng(4)
    g = np.zeros_like(E)
    idx = rng.choice(N, 10, replace=False)
    for c in range(3):
        g[3 * idx + c] = np.conj(O.local_field(al, P, diel)[3 * idx + c])
    m.set(T.VEC_E, g)
    m.set(T.VEC_P, np.zeros_like(E))
    m.solve(tol=1e-12)
    lam_ = m.get(T.VEC_P)[0]
    A = O.assemble_dense(R, 10.0, K, al)
    assert O.rel_l2(A.T @ lam_, g) < 1e-10
    # reciprocity: lambda^T E = g^T P
    assert abs(lam_ @ E - g @ P) < 1e-9 * abs(g @ P)


def test_batched_solves_equal_single_solves_bitwise(ctx):
    # right-hand sides with different polarisation / angle (different 1/alpha)
    Nx, Ny, Nz, d, lam = 10, 9, 6, 10.0, 500.0
    R = O.build_a_bulk(Nx, Ny, Nz)
    N = len(R) // 3
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = np.random.default_rng(2).random(N).repeat(3)
    ALs, Es = [], []
    for th in (0.0, 10.0, 20.0, 30.0):
        for pol in ("p", "y"):
            r = np.deg2rad(th)
            nK = np.array([np.sin(r), 0, np.cos(r)])
            nE = O.ne_perp_in_xz(r) if pol == "p" else np.array([0, 1.0, 0])
            ALs.append(O.diel_al(t, mat, lam, K, d, nE, nK)[1])
            Es.append(O.einc(R, K, d, 1.0, nK, nE))
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    mb = T.DeviceModel(core, 8)
    mb.set(T.VEC_AL, np.concatenate(ALs))
    mb.set(T.VEC_E, np.concatenate(Es))
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        mb.set(T.VEC_P, np.zeros(8 * 3 * N, complex))
        infos = mb.solve(mode=mode, tol=1e-8)
        Pb = mb.get(T.VEC_P)
        for j in (0, 3, 7):
            m1 = T.DeviceModel(core, 1)
            m1.set(T.VEC_AL, ALs[j])
            m1.set(T.VEC_E, Es[j])
            m1.set(T.VEC_P, np.zeros(3 * N, complex))
            i1 = m1.solve(mode=mode, tol=1e-8)[0]
            assert i1["iterations"] == infos[j]["iterations"]
            assert np.array_equal(m1.get(T.VEC_P)[0], Pb[j])
