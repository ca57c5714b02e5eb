"""GPU parity of the FCD interaction tensor (AProductCore::FCD_inter,
AProductCore.cpp:625-695, selected by A_dic_generator, 560-571) against the
oracle's restatement, whose Si / Ci tables are scipy.special.sici at the
0.1-spaced nodes interpolated linearly (SiCi.cpp:27-46).  The library
computes the node values itself (series and continued fraction), so the
tensor agrees to rounding, not bit for bit: tolerance 1e-12 relative for
the tensor, 1e-11 relative L2 for products."""
import math

import numpy as np
import pytest

from oracle import oracle as O

import topodda as T

pytestmark = pytest.mark.gpu


def test_tensor_matches_oracle():
    # the C-ABI tensor (the same code the Green kernel runs)
    rng = np.random.default_rng(3)
    for lam, d in [(542.0, 10.0), (500.0, 25.0), (1200.0, 60.0)]:
        K = O.wavenumber(lam)
        pts = list(rng.integers(-40, 41, size=(40, 3)) * d) + [(d, 0, 0), (0, 0, 2 * d), (d, d, d)]
        for p in pts:
            got = T.green_tensor_fcd(K, d, *p)
            want = O.fcd_inter(K, d, *p)
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max() + 1e-300, (lam, d, p)
    assert np.all(T.green_tensor_fcd(K, d, 0.0, 0.0, 0.0) == 0)


def test_tensor_rejects_coarse_lattice():
    K = O.wavenumber(100.0)
    with pytest.raises(T.TddaError):
        T.green_tensor_fcd(K, (math.pi - 0.05) / K, 1.0, 0.0, 0.0)


@pytest.mark.parametrize("shape,lam,d", [((7, 6, 5), 542.0, 10.0), ((9, 4, 6), 500.0, 25.0)])
def test_fcd_aproduct_full_box(ctx, shape, lam, d):
    Nx, Ny, Nz = shape
    print("FCD Aproduct", shape, lam, d, flush=True)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, method="FCD")
    R = O.build_a_bulk(Nx, Ny, Nz)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(3 * core.N) + 1j * rng.standard_normal(3 * core.N)
    y = core.Aproduct(b)
    assert O.rel_l2(y, O.aproduct_fcd_direct(R, d, core.K, b)) < 1e-11
    core.close()


def test_fcd_aproduct_sparse_sites(ctx):
    rng = np.random.default_rng(5)
    Nx, Ny, Nz, d, lam = 8, 7, 5, 10.0, 542.0
    R = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
    R = R[rng.random(len(R)) < 0.6].reshape(-1)
    print("FCD Aproduct sparse", len(R) // 3, flush=True)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam, method="FCD")
    b = rng.standard_normal(3 * core.N) + 1j * rng.standard_normal(3 * core.N)
    assert O.rel_l2(core.Aproduct(b), O.aproduct_fcd_direct(R, d, core.K, b)) < 1e-11
    core.close()


def test_fcd_periodic(ctx):
    Nx, Ny, Nz, d, lam = 6, 5, 3, 10.0, 542.0
    images = (1, 2, 65.0, 55.0)
    print("FCD periodic", images, flush=True)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, periodic=images, method="FCD")
    R = O.build_a_bulk(Nx, Ny, Nz)
    rng = np.random.default_rng(8)
    b = rng.standard_normal(3 * core.N) + 1j * rng.standard_normal(3 * core.N)
    assert O.rel_l2(core.Aproduct(b), O.aproduct_periodic_direct(R, d, core.K, b, *images, fcd=True)) < 1e-11
    core.close()


def test_fcd_rejects_coarse_lattice(ctx):
    # K d >= pi - 0.1 has Kf - K below the tables' first node
    lam = 2 * math.pi * 10.0 / (math.pi - 0.05)
    with pytest.raises(T.TddaError):
        T.AProductCore(ctx, None, 4, 4, 4, 10.0, lam, method="FCD")
    with pytest.raises(T.TddaError):
        T.AProductCore(ctx, None, 4, 4, 4, 10.0, 542.0, method="XYZ")


def test_fcd_solve_both_modes(ctx):
    # forward solves under the FCD operator in FP64 and mixed mode reach the
    # true-residual bar against the oracle's direct sum
    Nx, Ny, Nz, d, lam = 8, 7, 5, 10.0, 542.0
    print("FCD solve", (Nx, Ny, Nz), flush=True)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, method="FCD")
    rng = np.random.default_rng(21)
    K = core.K
    R = O.build_a_bulk(Nx, Ny, Nz)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = rng.random(len(R) // 3).repeat(3)
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
    E = O.einc(R, K, d, 1.0, nK, nE)
    m = T.DeviceModel(core, 1)
    m.set(T.VEC_AL, al)
    m.set(T.VEC_E, E)
    for mode in (T.MODE_FP64, T.MODE_MIXED):
        m.set(T.VEC_P, np.zeros_like(E))
        info = m.solve(mode=mode, tol=1e-8, true_residual=1)[0]
        assert info["status"] == T.CONVERGED and info["true_residual"] <= 1.0001e-8
        P = m.get(T.VEC_P)[0]
        assert O.rel_l2(O.aproduct_fcd_direct(R, d, K, P) + al * P, E) <= 1.0001e-8
    m.close()
    core.close()
