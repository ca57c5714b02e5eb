"""GPU parity of the field at points (DDAModel::solve_E, DDAModel.cpp:530-565,
tdda_field_at_points) and of PointE under the FCD tensor against the
oracle's direct sums.  The oracle's incident field keeps the reference's
cosine argument of components y and z (DDAModel.cpp:543-544); with nK_x = 0
that argument is the plane wave's and the field equals PointE's field at the
same point (ObjDDAModel.cpp:44-47), which pins the sum itself.
Tolerance: 1e-12 relative per point (FP64 reductions in another order)."""
import numpy as np
import pytest

from oracle import oracle as O
import topodda as T

pytestmark = pytest.mark.gpu


def _model(ctx, R, dims, d, lam, method="LDR", seed=0):
    Nx, Ny, Nz = dims
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam, method=method)
    m = T.DeviceModel(core, 1)
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(3 * core.N) + 1j * rng.standard_normal(3 * core.N)
    m.set(T.VEC_P, P)
    return core, m, P


@pytest.mark.parametrize("method", ["LDR", "FCD"])
@pytest.mark.parametrize("sparse", [False, True])
def test_field_at_points_matches_direct_sum(ctx, method, sparse):
    dims, d, lam = (6, 5, 4), 10.0, 542.0
    R = O.build_a_bulk(*dims)
    if sparse:
        keep = np.random.default_rng(2).random(len(R) // 3) < 0.7
        R = R.reshape(-1, 3)[keep].reshape(-1)
    print("field at points", method, "sparse" if sparse else "full", flush=True)
    core, m, P = _model(ctx, None if not sparse else R, dims, d, lam, method)
    # outside, on sites (own dipole excluded) and between them
    pts = np.array([[-3, 0, 0], [8, 2, 1], [0, 0, 0], [2, 3, 1], [5, 4, 3], [3, -4, 9], [1, 1, -2]])
    for nK in ([0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.48, 0.6, 0.64]):
        nE = [0.8, 0.0, -0.6] if nK[0] else [1.0, 0.0, 0.0]
        got = T.field_at_points(m, 0, pts, nK, nE, 1.3)
        want = O.field_at_points(R, d, core.K, P, pts, nK, nE, 1.3, fcd=method == "FCD")
        err = np.abs(got - want).max(axis=1) / np.abs(want).max(axis=1)
        assert err.max() < 1e-12, (nK, err)
        if nK[0] == 0.0:
            for q, p in enumerate(pts):
                Ep = O.point_e_field(R, d, core.K, P, d * p, nK, nE, 1.3, fcd=method == "FCD")[0]
                assert np.abs(got[q] - Ep).max() < 1e-12 * np.abs(Ep).max()
    m.close()
    core.close()


def test_field_at_points_edge_cases(ctx):
    core, m, P = _model(ctx, None, (3, 3, 2), 10.0, 542.0)
    assert T.field_at_points(m, 0, np.zeros((0, 3), np.int32), [0, 0, 1], [1, 0, 0]).shape == (0, 3)
    with pytest.raises(T.TddaError):
        T.field_at_points(m, 1, [[0, 0, 0]], [0, 0, 1], [1, 0, 0])
    m.close()
    core.close()


def test_point_e_uses_the_fcd_tensor(ctx):
    dims, d, lam = (5, 4, 3), 10.0, 542.0
    core, m, P = _model(ctx, None, dims, d, lam, "FCD", seed=4)
    R = O.build_a_bulk(*dims)
    r0, nK, nE = np.array([25.0, 18.0, 52.0]), np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    src = T.DeviceModel(core, 1)
    F = T.objective_point_e(m, 0, r0, nK, nE, 1.0, source=src)
    Ef, As = O.point_e_field(R, d, core.K, P, r0, nK, nE, 1.0, fcd=True)
    assert abs(F - np.linalg.norm(Ef)) < 1e-12 * abs(F)
    want = np.concatenate([-(A.T @ np.conj(Ef)) / F for A in As])
    got = src.get(T.VEC_E)[0]
    assert O.rel_l2(got, want) < 1e-12
    src.close()
    m.close()
    core.close()
