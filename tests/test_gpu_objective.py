"""GPU parity of the objectives and their closed-form adjoint sources
(tdda_objective_*; ObjDDAModel.cpp:20-177, EvoDDAModel.cpp:353-391) against
the oracle (tests/test_objective_oracle.py pins it by differentiation).

Tolerances: objective values 1e-12 relative; sources 1e-12 relative L2."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAT = [1.0, 2.25 + 0.01j]


def solved_model(ctx, R, Nx, Ny, Nz, d, lam, t, nrhs=2):
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    K = O.wavenumber(lam)
    Rf = O.build_a_bulk(Nx, Ny, Nz) if R is None else R
    mat = O.material_rule(np.array(MAT))
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
    E = O.einc(Rf, K, d, 1.0, nK, nE)
    m = T.DeviceModel(core, nrhs)
    m.set(T.VEC_AL, np.tile(al, nrhs))
    m.set(T.VEC_E, np.tile(E, nrhs))
    m.set(T.VEC_P, np.zeros(nrhs * len(E), complex))
    m.solve(tol=1e-10)
    return core, m, Rf, K, al, E, nK, nE


@pytest.mark.parametrize("sparse", [False, True])
def test_point_e_value_and_source(ctx, sparse):
    Nx, Ny, Nz, d, lam = 8, 7, 4, 10.0, 542.0
    print(f"PointE sparse={sparse}", flush=True)
    R = None
    if sparse:
        full = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
        keep = np.random.default_rng(2).random(len(full)) < 0.7
        R = full[keep].reshape(-1).astype(np.int32)
    N = Nx * Ny * Nz if R is None else len(R) // 3
    t = np.random.default_rng(1).random(N).repeat(3)
    core, m, Rf, K, al, E, nK, nE = solved_model(ctx, R, Nx, Ny, Nz, d, lam, t)
    P = m.get(T.VEC_P)[0]
    r0 = np.array([35.0, 30.0, 80.0])
    F = T.objective_point_e(m, 0, r0, nK, nE, 1.0, source=m, source_rhs=1)
    Fo = O.obj_point_e(Rf, d, K, P, r0, nK, nE, 1.0)
    assert abs(F - Fo) <= 1e-12 * Fo
    src = m.get(T.VEC_E)[1]
    assert O.rel_l2(src, O.devp_point_e(Rf, d, K, P, r0, nK, nE, 1.0)) < 1e-12
    assert np.array_equal(m.get(T.VEC_E)[0], E)  # rhs 0 untouched
    m.close()
    core.close()


def test_integrated_e_value_and_source(ctx):
    Nx, Ny, Nz, d, lam = 9, 6, 3, 10.0, 500.0
    t = np.random.default_rng(4).random(Nx * Ny * Nz).repeat(3)
    core, m, Rf, K, al, E, nK, nE = solved_model(ctx, None, Nx, Ny, Nz, d, lam, t)
    P = m.get(T.VEC_P)[0]
    for pw, box, eta, beta in [(2, (0, 5, 0, 5, 0, 5), 0.5, 0.0), (1, (1, 0, 0, 9, 0, 9), 0.4, 3.0),
                               (3, (0, 2, 0, 2, 0, 2), 0.5, 8.0)]:
        F = T.objective_integrated_e(m, 0, pw, box, eta, beta, t, source=m, source_rhs=1)
        Fo = O.obj_integrated_e(Rf, al, P, t, pw, box, eta, beta)
        assert abs(F - Fo) <= 1e-12 * abs(Fo)
        src = m.get(T.VEC_E)[1]
        assert O.rel_l2(src, O.devp_integrated_e(Rf, al, P, t, pw, box, eta, beta)) < 1e-12
    m.close()
    core.close()


def test_adjoint_solve_from_device_source(ctx):
    # EvoDDAModel.cpp:596-608 on the device: objective -> source in the
    # adjoint model's E -> solve; equals the solve with the same source set
    # from the host, and lambda^T E = devp^T P (A is complex symmetric)
    Nx, Ny, Nz, d, lam = 8, 8, 3, 10.0, 542.0
    t = np.random.default_rng(6).random(Nx * Ny * Nz).repeat(3)
    core, m, Rf, K, al, E, nK, nE = solved_model(ctx, None, Nx, Ny, Nz, d, lam, t, nrhs=1)
    P = m.get(T.VEC_P)[0]
    adj = T.DeviceModel(core, 1)
    adj.set(T.VEC_AL, al)
    r0 = np.array([40.0, 35.0, 60.0])
    T.objective_point_e(m, 0, r0, nK, nE, 1.0, source=adj, source_rhs=0)
    src = adj.get(T.VEC_E)[0]
    adj.set(T.VEC_P, np.zeros_like(E))
    i1 = adj.solve(tol=1e-10)[0]
    lam1 = adj.get(T.VEC_P)[0]
    host = T.DeviceModel(core, 1)
    host.set(T.VEC_AL, al)
    host.set(T.VEC_E, src)
    host.set(T.VEC_P, np.zeros_like(E))
    i2 = host.solve(tol=1e-10)[0]
    assert i1["iterations"] == i2["iterations"] and np.array_equal(lam1, host.get(T.VEC_P)[0])
    assert abs(lam1 @ E - src @ P) < 1e-8 * abs(src @ P)
    for h in (host, adj, m):
        h.close()
    core.close()


def test_objectives_reject_bad_input(ctx):
    core = T.AProductCore(ctx, None, 3, 3, 2, 10.0, 500.0)
    other = T.AProductCore(ctx, None, 3, 3, 2, 10.0, 500.0)
    m = T.DeviceModel(core, 1)
    o = T.DeviceModel(other, 1)
    z = np.zeros(3)
    with pytest.raises(T.TddaError):
        T.objective_point_e(m, 1, z, [0, 0, 1.0], [1.0, 0, 0])                # rhs out of range
    with pytest.raises(T.TddaError):
        T.objective_point_e(m, 0, z, [0, 0, 1.0], [1.0, 0, 0], source=o)      # other operator
    with pytest.raises(T.TddaError):
        T.objective_integrated_e(m, 0, 0, (0, 5, 0, 5, 0, 5), 0.5, 1.0, np.zeros(54))  # pow < 1
    with pytest.raises(T.TddaError):
        T.objective_integrated_e(m, 0, 2, (0, 5, 0, 5, 0, 5), 0.5, 1.0, np.full(54, 2.0))  # diel_old > 1
    for h in (m, o, core, other):
        h.close()
