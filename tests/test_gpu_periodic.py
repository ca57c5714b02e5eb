"""GPU parity of the periodic operator (tdda_op_create_periodic; the
reference's periodic AProductCore, AProductCore.cpp:251-536, 573-586)
against the oracle's direct sum over the image lattice
(tests/test_periodic_oracle.py pins that sum).

Tolerances: complex128 <= 1e-12 relative L2; complex64 <= 5e-6."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,images", [((4, 3, 2), (1, 1, 40.0, 30.0)), ((5, 5, 1), (3, 2, 50.0, 50.0)),
                                         ((3, 4, 3), (2, 0, 45.0, 70.0))])
def test_periodic_aproduct_matches_image_sum(ctx, dims, images):
    Nx, Ny, Nz = dims
    d, lam = 10.0, 500.0
    print(f"periodic aproduct {dims} images {images}", flush=True)
    R = O.build_a_bulk(Nx, Ny, Nz)
    rng = np.random.default_rng(Nx + Ny)
    b = rng.standard_normal(len(R)) + 1j * rng.standard_normal(len(R))
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, periodic=images)
    ref = O.aproduct_periodic_direct(R, d, O.wavenumber(lam), b, *images)
    assert O.rel_l2(core.Aproduct(b), ref) < 1e-12
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < 5e-6
    core.close()


def test_periodic_solve_and_symmetry(ctx):
    # an 8 x 8 x 2 unit cell with 5 x 5 images: complex-symmetric operator,
    # forward solves converge in both modes to the true-residual bar
    Nx, Ny, Nz, d, lam = 8, 8, 2, 10.0, 542.0
    images = (2, 2, 80.0, 80.0)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, periodic=images)
    rng = np.random.default_rng(4)
    N3 = 3 * Nx * Ny * Nz
    b = rng.standard_normal(N3) + 1j * rng.standard_normal(N3)
    c = rng.standard_normal(N3) + 1j * rng.standard_normal(N3)
    Ab, Ac = core.Aproduct(b), core.Aproduct(c)
    assert abs(c @ Ab - b @ Ac) < 1e-12 * abs(c @ Ab)
    K = O.wavenumber(lam)
    R = O.build_a_bulk(Nx, Ny, Nz)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = rng.random(len(R) // 3).repeat(3)
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
    E = O.einc(R, K, d, 1.0, nK, nE)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    ref = None
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        m.set(T.VEC_P, np.zeros_like(E))
        info = m.solve(mode=mode, tol=1e-8, true_residual=1)[0]
        assert info["status"] == T.CONVERGED and info["true_residual"] <= 1.0001e-8
        P = m.get(T.VEC_P)[0]
        # residual against the oracle's image sum
        y = O.aproduct_periodic_direct(R, d, K, P, *images) + al * P
        assert O.rel_l2(y, E) <= 1.0001e-8
        ref = P if ref is None else ref
        assert O.rel_l2(P, ref) < 1e-6
    m.close()
    core.close()


def test_periodic_rejects_bad_images(ctx):
    for images in [(-1, 0, 10.0, 10.0), (1, 1, 0.0, 10.0), (1, 1, 10.0, float("nan")), (1000, 1000, 10.0, 10.0)]:
        with pytest.raises(T.TddaError):
            T.AProductCore(ctx, None, 3, 3, 2, 10.0, 500.0, periodic=images)
