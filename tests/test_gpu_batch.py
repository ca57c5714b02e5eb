"""GPU tests of batches: right-hand sides of several structures (different
boxes, FFT grids, wavelengths and incidences) solved by one set of launches
per pass.  Each result must equal, bit for bit, the same right-hand side
solved alone (the reference solves one DDAModel at a time:
EvoDDAModel.cpp:433-441, DDAModel.cpp:131-260), and the converged solutions
must match the oracle."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (Nx, Ny, Nz, d, lam, nrhs): grids 20x18x12 (runtime-plan x, compiled y/z
# mix), 24x15x10, 32x32x16, 16x12x9 -- no two share all three lengths
STRUCTS = [(10, 9, 6, 10.0, 500.0, 2), (12, 7, 5, 10.0, 542.0, 1), (16, 16, 8, 8.0, 600.0, 3),
           (8, 6, 5, 10.0, 450.0, 2)]


def inputs(Nx, Ny, Nz, d, lam, nrhs, seed):
    """Per right-hand side (1/alpha, E): incidence angle and polarisation
    vary, so 1/alpha differs per right-hand side (tools.cpp:617-630)."""
    R = O.build_a_bulk(Nx, Ny, Nz)
    N = len(R) // 3
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = np.random.default_rng(seed).random(N).repeat(3)
    al, E = [], []
    for j in range(nrhs):
        th = np.deg2rad(10.0 * j)
        nK = np.array([np.sin(th), 0.0, np.cos(th)])
        nE = O.ne_perp_in_xz(th) if j % 2 == 0 else np.array([0.0, 1.0, 0.0])
        al.append(O.diel_al(t, mat, lam, K, d, nE, nK)[1])
        E.append(O.einc(R, K, d, 1.0, nK, nE))
    return R, N, K, np.array(al), np.array(E)


def make_models(ctx, structs=STRUCTS):
    out = []
    for i, (Nx, Ny, Nz, d, lam, nrhs) in enumerate(structs):
        R, N, K, al, E = inputs(Nx, Ny, Nz, d, lam, nrhs, seed=10 + i)
        core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
        m = T.DeviceModel(core, nrhs)
        m.set(T.VEC_AL, al)
        m.set(T.VEC_E, E)
        m.set(T.VEC_P, np.zeros_like(E))
        out.append((core, m, R, K, al, E))
    return out


@pytest.mark.parametrize("mode", [T.MODE_FP64, T.MODE_MIXED])
def test_batch_solve_equals_single_solves_bitwise(ctx, mode):
    print(f"batch solve mode {mode}", flush=True)
    models = make_models(ctx)
    grids = [c.grid for c, *_ in models]
    assert len(set(grids)) == len(grids)
    b = T.Batch([m for _, m, *_ in models])
    assert b.size == sum(s[5] for s in STRUCTS)
    infos = b.solve(mode=mode, tol=1e-8)
    k = 0
    for (core, m, R, K, al, E) in models:
        Pb = m.get(T.VEC_P)
        alone = T.DeviceModel(core, 1)
        for j in range(m.nrhs):
            alone.set(T.VEC_AL, al[j])
            alone.set(T.VEC_E, E[j])
            alone.set(T.VEC_P, np.zeros_like(E[j]))
            i1 = alone.solve(mode=mode, tol=1e-8)[0]
            assert infos[k]["status"] == T.CONVERGED == i1["status"]
            assert infos[k]["iterations"] == i1["iterations"]
            assert infos[k]["error"] == i1["error"]
            assert np.array_equal(alone.get(T.VEC_P)[0], Pb[j])
            k += 1
        alone.close()
    b.close()


def test_batch_solutions_match_oracle(ctx):
    models = make_models(ctx, STRUCTS[:2])
    b = T.Batch([m for _, m, *_ in models])
    infos = b.solve(tol=1e-8, true_residual=1)
    k = 0
    for (core, m, R, K, al, E), s in zip(models, STRUCTS):
        Nx, Ny, Nz, d = s[:4]
        op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
        P = m.get(T.VEC_P)
        for j in range(m.nrhs):
            print(f"oracle check struct {s} rhs {j}", flush=True)
            assert infos[k]["status"] == T.CONVERGED and infos[k]["true_residual"] <= 1.0001e-8
            Pref, st, _, _ = op.bicgstab(al[j], E[j], max_it=5000, tol=1e-13)
            assert st == 0
            assert O.rel_l2(P[j], Pref) < 1e-6
            assert O.rel_l2(op.apply(P[j], al[j]), E[j]) <= 1.0001e-8
            k += 1
        op.close()


@pytest.mark.parametrize("prec", [T.C128, T.C64])
def test_batch_apply_equals_model_apply(ctx, prec):
    models = make_models(ctx)
    rng = np.random.default_rng(7)
    xs = []
    for (_, m, *_rest) in models:
        x = rng.standard_normal((m.nrhs, m.n3)) + 1j * rng.standard_normal((m.nrhs, m.n3))
        m.set(T.VEC_P, x)
        xs.append(x)
    b = T.Batch([m for _, m, *_ in models])
    b.apply(prec)
    yb = [m.get(T.VEC_Y) for _, m, *_ in models]
    for (_, m, *_rest), y in zip(models, yb):
        assert np.array_equal(m.apply(prec), y)
    b.close()


def test_batch_keeps_models_alive_and_rejects_bad_input(ctx):
    models = make_models(ctx, STRUCTS[:2])
    b = T.Batch([m for _, m, *_ in models])
    # the batch holds its own references: closing the models keeps their state
    for _, m, *_ in models:
        m.close()
    infos = b.solve(tol=1e-6)
    assert all(i["status"] == T.CONVERGED for i in infos)
    b.close()
    with pytest.raises(T.TddaError):
        T.Batch([])
    other = T.Context(0)
    try:
        c2 = T.AProductCore(other, O.build_a_bulk(4, 4, 4), 4, 4, 4, 10.0, 500.0)
        m2 = T.DeviceModel(c2, 1)
        c1 = T.AProductCore(ctx, O.build_a_bulk(4, 4, 4), 4, 4, 4, 10.0, 500.0)
        m1 = T.DeviceModel(c1, 1)
        with pytest.raises(T.TddaError):
            T.Batch([m1, m2])
        m1.close()
        m2.close()
        c1.close()
        c2.close()
    finally:
        other.close()
    # complex64-only models cannot join a batch (solves need complex128)
    c3 = T.AProductCore(ctx, O.build_a_bulk(4, 4, 4), 4, 4, 4, 10.0, 500.0, precisions=T.C64)
    m3 = T.DeviceModel(c3, 1)
    with pytest.raises(T.TddaError):
        T.Batch([m3])
    m3.close()
    c3.close()
