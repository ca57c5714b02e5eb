"""GPU tests of right Jacobi preconditioning (tdda_solve_opts.precond =
TDDA_PRECOND_JACOBI, no counterpart in the reference): BiCGSTAB
(DDAModel.cpp:131-260) on A D^-1 with D = diag(al), P = D^-1 y.  The
preconditioned solves must reach the same 1e-8 true residual, checked with the
oracle operator (AProductCore.cpp:697-758 + the al o P term of
DDAModel.cpp:577-586), and the same solution as the unpreconditioned solver
within what two 1e-8 residuals allow; a batch equals its members alone bit for
bit, and the request loop's preconditioned results match the model's."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def problem(Nx, Ny, Nz, seed, frac=1.0, d=10.0, lam=500.0):
    rng = np.random.default_rng(seed)
    R = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
    if frac < 1:
        R = R[rng.random(len(R)) < frac]
    R = R.reshape(-1).astype(np.int32)
    N = len(R) // 3
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = rng.random(N).repeat(3)  # design values: eps between 1.01 and 2.25 + 0.01i
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
    E = O.einc(R, K, d, 1.0, nK, nE)
    return R, N, K, al, E


# fused z + y passes (24 x 20 x 10), separate passes (16 x 12 x 30: z length
# 60), sparse sites through the cell map
@pytest.mark.parametrize("dims,frac", [((24, 20, 10), 1.0), ((16, 12, 30), 1.0), ((20, 18, 9), 0.6)])
def test_jacobi_solves_match_unpreconditioned(ctx, dims, frac):
    Nx, Ny, Nz = dims
    d, lam = 10.0, 500.0
    print(f"jacobi {dims} frac={frac}", flush=True)
    R, N, K, al, E = problem(Nx, Ny, Nz, seed=Nx, frac=frac)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    out = {}
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        for pc in (T.PRECOND_NONE, T.PRECOND_JACOBI):
            info = m.solve(mode=mode, tol=1e-8, x0_zero=1, true_residual=1, precond=pc)[0]
            x = m.get(T.VEC_P)[0]
            assert info["status"] == T.CONVERGED, (mode, pc, info)
            res = O.rel_l2(op.apply(x) + al * x, E)
            assert res <= 1e-8, (mode, pc, res)
            assert abs(info["true_residual"] - res) <= 1e-9
            out[(mode, pc)] = (x, info["iterations"])
    x_ref, it_ref = out[(T.MODE_FP64, T.PRECOND_NONE)]
    for key, (x, it) in out.items():
        assert O.rel_l2(x, x_ref) < 1e-6, key
    # design values spread 1/alpha over two orders of magnitude: the
    # diagonal scaling removes most of the iterations
    assert out[(T.MODE_FP64, T.PRECOND_JACOBI)][1] * 3 <= it_ref
    m.close()
    core.close()


def test_jacobi_from_a_given_start(ctx):
    # x0 given: r = E - A x0, then y0 = D x0; from a converged P the solve
    # stops at once with the same P
    Nx, Ny, Nz = 24, 20, 10
    R, N, K, al, E = problem(Nx, Ny, Nz, seed=3)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0)
    op = O.OrcOperator(R, Nx, Ny, Nz, 10.0, K, grid=core.grid)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    m.solve(mode=T.MODE_FP64, tol=1e-10, x0_zero=1, precond=T.PRECOND_JACOBI)
    x1 = m.get(T.VEC_P)[0]
    info = m.solve(mode=T.MODE_FP64, tol=1e-8, x0_zero=0, true_residual=1, precond=T.PRECOND_JACOBI)[0]
    x2 = m.get(T.VEC_P)[0]
    assert info["status"] == T.CONVERGED and info["iterations"] <= 1, info
    assert O.rel_l2(x2, x1) < 1e-9
    assert O.rel_l2(op.apply(x2) + al * x2, E) <= 1e-8
    # a half-converged start continues to the same solution
    m.set(T.VEC_P, 0.5 * x1)
    info = m.solve(mode=T.MODE_FP64, tol=1e-8, x0_zero=0, true_residual=1, precond=T.PRECOND_JACOBI)[0]
    x3 = m.get(T.VEC_P)[0]
    assert info["status"] == T.CONVERGED
    assert O.rel_l2(op.apply(x3) + al * x3, E) <= 1e-8
    m.close()
    core.close()


def test_jacobi_batch_equals_alone_bitwise(ctx):
    shapes = [(24, 20, 10), (16, 12, 30), (8, 10, 6)]
    models, solo = [], []
    for k, (Nx, Ny, Nz) in enumerate(shapes):
        R, N, K, al, E = problem(Nx, Ny, Nz, seed=10 + k)
        core = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0)
        for lst in (models, solo):
            m = T.DeviceModel(core, 2)
            m.set(T.VEC_AL, np.concatenate([al, al]))
            m.set(T.VEC_E, np.concatenate([E, 1j * E]))
            lst.append(m)
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        b = T.Batch(models)
        binfo = b.solve(mode=mode, tol=1e-8, x0_zero=1, true_residual=1, precond=T.PRECOND_JACOBI)
        b.close()
        k = 0
        for m, s in zip(models, solo):
            sinfo = s.solve(mode=mode, tol=1e-8, x0_zero=1, true_residual=1, precond=T.PRECOND_JACOBI)
            for j in range(2):
                assert binfo[k]["iterations"] == sinfo[j]["iterations"]
                assert binfo[k]["true_residual"] <= 1e-8
                k += 1
            assert np.array_equal(m.get(T.VEC_P), s.get(T.VEC_P)), mode
    for m in models + solo:
        m.close()


def test_unknown_preconditioner_is_an_error(ctx):
    R, N, K, al, E = problem(8, 8, 6, seed=1)
    core = T.AProductCore(ctx, R, 8, 8, 6, 10.0, 500.0)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    with pytest.raises(T.TddaError):
        m.solve(mode=T.MODE_FP64, tol=1e-8, x0_zero=1, precond=7)
    m.close()
    core.close()


@pytest.mark.parametrize("mode", [T.MODE_MIXED, T.MODE_FP64])
def test_server_jacobi_matches_model_solves(ctx, mode):
    from test_gpu_serve import SPECS, model_solve, submit
    srv = T.Server(0, budget_bytes=8e9, mode=mode, tol=1e-8, precond=T.PRECOND_JACOBI)
    tickets = [submit(srv, s, i) for i, s in enumerate(SPECS)]
    for i, (s, tk) in enumerate(zip(SPECS, tickets)):
        print(f"server jacobi request {s} mode {mode}", flush=True)
        assert srv.wait(tk, 120.0) == 1
        P, infos, _ = srv.result(tk)
        Pm, _, al, E, R, K = model_solve(ctx, s, i, mode)
        op = O.OrcOperator(R, s[0], s[1], s[2], s[3], K)
        for j in range(s[5]):
            assert infos[j]["status"] == T.CONVERGED
            assert infos[j]["true_residual"] <= 1.0001e-8
            assert O.rel_l2(P[j], Pm[j]) < 1e-6
            assert O.rel_l2(op.apply(P[j], al[j]), E[j]) <= 1.0001e-8
        op.close()
    st = srv.stats()
    assert st["completed"] == len(SPECS) and st["failed"] == 0
    srv.close()
