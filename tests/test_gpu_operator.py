"""GPU parity of the operator (AProductCore::Aproduct, AProductCore.cpp:697-758,
and DDAModel::Aproductwithalb, DDAModel.cpp:577-586) against the oracle.

Tolerances: complex128 path <= 1e-12 relative L2 (the FP64 bar for this
path); complex64 path <= 5e-6 relative L2 (single-precision FFT rounding).
"""
import os

import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL64 = 1e-12
TOL32 = 5e-6


def problem(Nx, Ny, Nz, d, lam, seed=0, frac=1.0):
    rng = np.random.default_rng(seed)
    R = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
    if frac < 1:
        R = R[rng.random(len(R)) < frac]
    R = R.reshape(-1).astype(np.int32)
    N = len(R) // 3
    b = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    al = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    return R, N, b, al


def test_aproduct_matches_long_double_fixture(ctx):
    g = np.load(os.path.join(GOLD, "aproduct_tiny.npz"), allow_pickle=False)
    for name in ("a", "b", "c"):
        Nx, Ny, Nz = (int(v) for v in g[f"{name}_dims"])
        print(f"fixture {name} {Nx}x{Ny}x{Nz}", flush=True)
        core = T.AProductCore(ctx, g[f"{name}_R"], Nx, Ny, Nz, float(g[f"{name}_d"]), float(g[f"{name}_lam"]))
        y = g[f"{name}_y"]
        assert O.rel_l2(core.Aproduct(g[f"{name}_b"]), y) < TOL64
        assert O.rel_l2(core.Aproduct(g[f"{name}_b"], precision=T.C64), y) < TOL32


@pytest.mark.parametrize("dims,d,lam,frac", [((16, 16, 16), 10.0, 500.0, 1.0),
                                             ((20, 12, 9), 10.0, 542.0, 0.6),
                                             ((1, 7, 3), 10.0, 500.0, 1.0),
                                             ((9, 1, 1), 5.0, 450.0, 1.0)])
def test_aproduct_matches_oracle(ctx, dims, d, lam, frac):
    Nx, Ny, Nz = dims
    print(f"aproduct {dims} frac={frac}", flush=True)
    R, N, b, al = problem(Nx, Ny, Nz, d, lam, seed=Nx, frac=frac)
    K = O.wavenumber(lam)
    core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam)
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=core.grid)
    ref = op.apply(b)
    assert O.rel_l2(core.Aproduct(b), ref) < TOL64
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < TOL32
    assert O.rel_l2(core.Aproduct(b, al=al), ref + al * b) < TOL64


def test_aproduct_config2_size(ctx):
    # 80 x 80 x 16 slab, d = 10 nm, lambda = 542 nm, grid 160 x 160 x 32
    print("aproduct 80x80x16", flush=True)
    R, N, b, al = problem(80, 80, 16, 10.0, 542.0, seed=2)
    core = T.AProductCore(ctx, R, 80, 80, 16, 10.0, 542.0)
    assert core.grid == (160, 160, 32)
    ref = O.aproduct_numpy_fft(R, 80, 80, 16, 10.0, O.wavenumber(542.0), b, grid=core.grid)
    assert O.rel_l2(core.Aproduct(b), ref) < TOL64


@pytest.mark.parametrize("dims", [(512, 3, 2), (300, 4, 3)])
def test_aproduct_long_x_pencils(ctx, dims):
    # Mx = 1024 / 600: the complex128 x pass holds 3 columns of Mx points
    # (more than 64 KB of LDS for 1024); full box given as R = NULL
    Nx, Ny, Nz = dims
    print(f"aproduct long x {dims}", flush=True)
    R, N, b, al = problem(Nx, Ny, Nz, 10.0, 500.0, seed=Nx)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, 10.0, 500.0)
    assert core.grid[0] == (1024 if Nx == 512 else 600)
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, 10.0, O.wavenumber(500.0), b, grid=core.grid)
    assert O.rel_l2(core.Aproduct(b), ref) < TOL64
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < TOL32


@pytest.mark.parametrize("dims,grid", [((80, 80, 16), (160, 160, 32)),
                                       ((128, 20, 6), (256, 40, 12)),
                                       ((96, 96, 4), (192, 192, 8)),
                                       ((10, 10, 64), (20, 20, 128)),
                                       ((100, 100, 8), (320, 240, 16))])
def test_aproduct_inplace_tiles(ctx, dims, grid):
    # in-place LDS transforms: x tiles of 12 pencils (160; 256 in complex64)
    # or 6 (192, 256, 320 in complex128), y tiles of 128-byte rows (160 -
    # 256), z tiles (32, 128), also on a grid larger than the automatic one
    Nx, Ny, Nz = dims
    print(f"aproduct in-place tiles {dims} grid {grid}", flush=True)
    R, N, b, al = problem(Nx, Ny, Nz, 10.0, 500.0, seed=Ny)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    assert core.grid == grid
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, 10.0, O.wavenumber(500.0), b, grid=grid)
    assert O.rel_l2(core.Aproduct(b), ref) < TOL64
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < TOL32
    core.close()


def test_grid_size_independence(ctx):
    # any FFT grid >= 2N-1 gives the same linear convolution (the reference uses 2N-1)
    R, N, b, _ = problem(6, 5, 4, 10.0, 500.0, seed=4)
    ys = []
    for grid in [(0, 0, 0), (12, 10, 8), (15, 12, 9), (16, 18, 20)]:
        core = T.AProductCore(ctx, R, 6, 5, 4, 10.0, 500.0, grid=grid)
        ys.append(core.Aproduct(b))
    for y in ys[1:]:
        assert O.rel_l2(y, ys[0]) < 1e-13


def test_batched_aproduct_per_rhs_and_shared_diag(ctx):
    R, N, b, al = problem(7, 6, 5, 10.0, 500.0, seed=5)
    rng = np.random.default_rng(6)
    B = rng.standard_normal((3, 3 * N)) + 1j * rng.standard_normal((3, 3 * N))
    AL = rng.standard_normal((3, 3 * N)) + 1j * rng.standard_normal((3, 3 * N))
    core = T.AProductCore(ctx, R, 7, 6, 5, 10.0, 500.0)
    Y = core.Aproduct(B, al=AL)
    Ys = core.Aproduct(B, al=al)
    for j in range(3):
        yj = core.Aproduct(B[j])
        assert O.rel_l2(Y[j], yj + AL[j] * B[j]) < 1e-14
        assert O.rel_l2(Ys[j], yj + al * B[j]) < 1e-14


def test_sparse_geometry_with_padded_box(ctx):
    # grid header larger than the occupied extent (cf. the reference's sample
    # grid files, 83 vs 81): unoccupied cells are zero
    rng = np.random.default_rng(8)
    Rfull = O.build_a_bulk(9, 8, 5).reshape(-1, 3)
    keep = (Rfull[:, 0] < 7) & (Rfull[:, 1] < 6) & (rng.random(len(Rfull)) < 0.8)
    R = Rfull[keep]
    R = R[rng.permutation(len(R))].reshape(-1).astype(np.int32)  # arbitrary site order
    N = len(R) // 3
    b = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    core = T.AProductCore(ctx, R, 9, 8, 5, 10.0, 500.0)
    ref = O.aproduct_direct(R, 10.0, O.wavenumber(500.0), b)
    assert O.rel_l2(core.Aproduct(b), ref) < TOL64


def test_invalid_geometry_is_rejected(ctx):
    R = np.array([0, 0, 0, 0, 0, 0], np.int32)  # two sites in one cell
    with pytest.raises(T.TddaError):
        T.AProductCore(ctx, R, 2, 2, 2, 10.0, 500.0)
    with pytest.raises(T.TddaError):
        T.AProductCore(ctx, np.array([0, 0, 5], np.int32), 2, 2, 2, 10.0, 500.0)
    with pytest.raises(T.TddaError):
        T.AProductCore(ctx, O.build_a_bulk(4, 4, 4), 4, 4, 4, 10.0, 500.0, grid=(7, 7, 6))


def test_full_size_properties_config3(ctx):
    # 128 x 128 x 32 (grid 256 x 256 x 64): linearity, complex symmetry and
    # c64 / c128 agreement -- size-independent checks at the full size
    print("aproduct 128x128x32", flush=True)
    R, N, b, _ = problem(128, 128, 32, 8.0, 600.0, seed=3)
    rng = np.random.default_rng(33)
    c = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    core = T.AProductCore(ctx, R, 128, 128, 32, 8.0, 600.0)
    Y = core.Aproduct(np.stack([b, c, 2.0 * b - 3j * c]))
    assert O.rel_l2(Y[2], 2.0 * Y[0] - 3j * Y[1]) < 1e-13
    lhs, rhs = c @ Y[0], b @ Y[1]
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), Y[0]) < TOL32
