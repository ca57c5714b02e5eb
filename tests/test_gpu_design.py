"""GPU parity of the device-resident design update (tdda_design_*): the
reference's CoreStructure::UpdateStr with the cone filter and SmoothDensity
(CoreStructure.cpp:31-104, SpacePara.cpp:1241-1270, FilterOption.cpp:49-60)
followed by DDAModel::UpdateAlpha (DDAModel.cpp:470-484) for every
right-hand side, against the oracle (itself pinned to the reference's own
compiled classes, tests/test_design_oracle.py).

Tolerances: design values 1e-14 absolute (FP64 sums in the reference's order;
exp may differ by an ulp); 1/alpha 1e-13 relative."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAT = [1.0, 2.25 + 0.01j]


def incidences(n):
    out = []
    for j in range(n):
        th = np.deg2rad(12.0 * j)
        out.append((np.array([np.sin(th), 0.0, np.cos(th)]),
                    O.ne_perp_in_xz(th) if j % 2 == 0 else np.array([0.0, 1.0, 0.0])))
    return out


def oracle_al(para, Nx, Ny, Nz, lam, d, incs):
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array(MAT))
    t = O.design_diel_old(para, Nx, Ny, Nz)
    return np.array([O.diel_al(t, mat, lam, K, d, nE, nK)[1] for nK, nE in incs])


@pytest.mark.parametrize("dims,r,beta,eta", [((12, 10, 4), 2.0, 5.0, 0.5), ((20, 16, 3), 3.3, 0.0, 0.4)])
def test_design_steps_match_oracle(ctx, dims, r, beta, eta):
    Nx, Ny, Nz = dims
    d, lam = 10.0, 542.0
    print(f"design update {dims} r={r} beta={beta}", flush=True)
    incs = incidences(2)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
    m = T.DeviceModel(core, len(incs))
    rng = np.random.default_rng(Nx)
    p0 = rng.random(Nx * Ny)
    dz = T.Design(m, p0, r, O.material_rule(np.array(MAT)), incs)
    o_ref = p0
    for it in range(3):
        st = 0.2 * rng.standard_normal(Nx * Ny)
        dz.step(st, beta=beta, eta=eta)
        o_ref, f_ref, p_ref = O.design_update(o_ref, st, Nx, Ny, r, beta, eta)
        o, f, p = dz.get()
        assert np.array_equal(o, o_ref)
        assert np.max(np.abs(f - f_ref)) <= 1e-14
        assert np.max(np.abs(p - p_ref)) <= 1e-14
        al = m.get(T.VEC_AL)
        al_ref = oracle_al(p, Nx, Ny, Nz, lam, d, incs)
        for j in range(len(incs)):
            assert O.rel_l2(al[j], al_ref[j]) <= 1e-13
    dz.close()
    m.close()
    core.close()


def test_design_update_then_solve(ctx):
    # the updated 1/alpha (complex128 and complex64 copies) is what the
    # solvers use: solves after a device update equal solves with the same
    # 1/alpha set from the host
    Nx, Ny, Nz, d, lam = 10, 9, 4, 10.0, 500.0
    incs = incidences(2)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
    K = O.wavenumber(lam)
    R = O.build_a_bulk(Nx, Ny, Nz)
    E = np.concatenate([O.einc(R, K, d, 1.0, nK, nE) for nK, nE in incs])
    m = T.DeviceModel(core, 2)
    m.set(T.VEC_E, E)
    rng = np.random.default_rng(3)
    dz = T.Design(m, rng.random(Nx * Ny), 2.0, O.material_rule(np.array(MAT)), incs)
    dz.step(0.3 * rng.standard_normal(Nx * Ny), beta=2.0, eta=0.5)
    al = m.get(T.VEC_AL)
    h = T.DeviceModel(core, 2)
    h.set(T.VEC_E, E)
    h.set(T.VEC_AL, al.reshape(-1))
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        outs = []
        for mm in (m, h):
            mm.set(T.VEC_P, np.zeros_like(E))
            infos = mm.solve(mode=mode, tol=1e-8, true_residual=1)
            assert all(i["status"] == T.CONVERGED and i["true_residual"] <= 1.0001e-8 for i in infos)
            outs.append((mm.get(T.VEC_P), [i["iterations"] for i in infos]))
        assert outs[0][1] == outs[1][1]
        assert np.array_equal(outs[0][0], outs[1][0])
    dz.close()
    h.close()
    m.close()
    core.close()


def test_design_rejects_bad_input(ctx):
    core = T.AProductCore(ctx, None, 4, 4, 2, 10.0, 500.0)
    m = T.DeviceModel(core, 1)
    mat = O.material_rule(np.array(MAT))
    inc = incidences(1)
    with pytest.raises(T.TddaError):
        T.Design(m, np.full(16, 1.5), 2.0, mat, inc)        # outside [0, 1]
    with pytest.raises(T.TddaError):
        T.Design(m, np.full(16, 0.5), 0.05, mat, inc)       # no column in range
    with pytest.raises(T.TddaError):
        T.Design(m, np.full(16, 0.5), 2.0, mat[:1], inc)    # one material
    dz = T.Design(m, np.full(16, 0.5), 2.0, mat, inc)
    with pytest.raises(T.TddaError):
        dz.step(np.zeros(16), beta=1.0, eta=1.0)              # eta must lie in (0, 1)
    dz.close()
    # a sparse geometry has no full-box columns
    R = O.build_a_bulk(4, 4, 2).reshape(-1, 3)[:-3].reshape(-1).astype(np.int32)
    sparse = T.AProductCore(ctx, R, 4, 4, 2, 10.0, 500.0)
    ms = T.DeviceModel(sparse, 1)
    with pytest.raises(T.TddaError):
        T.Design(ms, np.full(16, 0.5), 2.0, mat, inc)
    ms.close()
    sparse.close()
    m.close()
    core.close()
