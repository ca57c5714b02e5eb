"""GPU parity of the register-resident x and y passes (k_pxr: four-step split
L = L1 * L2 with both sub-transforms in registers, the Green product between
them) against the numpy FFT convolution of the oracle and against the
LDS-staged x pass, on the lengths where the split takes each of its forms:
both factors radix kernels (32 = 4 x 8), composite factors (96 = 8 x 12,
180 = 12 x 15), odd lengths (45 = 5 x 9, 75 = 5 x 15, 81 = 9 x 9,
135 = 9 x 15, 225 = 15 x 15) and the complex128 kernels compiled for two
waves (192, 240, 256).  Nx is the largest the grid allows (Mx >= 2 Nx - 1)
or smaller, on full and sparse structures.

Tolerances: complex128 <= 1e-12 relative L2; complex64 <= 5e-6.
"""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
TOL32 = 5e-6

CASES = [  # (Nx, Ny, Nz), grid, occupied fraction
    ((16, 5, 3), (32, 10, 8), 1.0),
    ((23, 4, 3), (45, 8, 8), 0.6),
    ((38, 4, 2), (75, 8, 8), 1.0),
    ((41, 3, 3), (81, 8, 8), 0.7),
    ((40, 6, 4), (96, 12, 8), 1.0),
    ((68, 3, 2), (135, 8, 8), 1.0),
    ((80, 5, 3), (160, 10, 8), 0.5),
    ((90, 3, 2), (180, 8, 8), 1.0),
    ((96, 4, 2), (192, 8, 8), 0.8),
    ((113, 3, 2), (225, 8, 8), 1.0),
    ((100, 3, 3), (240, 8, 8), 1.0),
    ((128, 4, 2), (256, 8, 8), 0.5),
]


def problem(Nx, Ny, Nz, frac, seed):
    rng = np.random.default_rng(seed)
    R = O.build_a_bulk(Nx, Ny, Nz).reshape(-1, 3)
    if frac < 1:
        R = R[rng.random(len(R)) < frac]
    R = R.reshape(-1).astype(np.int32)
    N = len(R) // 3
    b = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    return R, b


@pytest.mark.parametrize("dims,grid,frac", CASES)
def test_register_x_pass(ctx, monkeypatch, dims, grid, frac):
    Nx, Ny, Nz = dims
    print(f"register x pass {dims} grid {grid} frac {frac}", flush=True)
    R, b = problem(Nx, Ny, Nz, frac, seed=grid[0])
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, 10.0, O.wavenumber(500.0), b, grid=grid)
    monkeypatch.setenv("TDDA_XPASS", "1")
    ys = []
    for nt in ("512", "256"):  # workgroups of 512 threads (wide tiles) and of 256
        monkeypatch.setenv("TDDA_XNT", nt)
        reg = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
        assert reg.grid == grid
        assert reg.xpass(T.C128) == "register" and reg.xpass(T.C64) == "register"
        y64 = reg.Aproduct(b)
        y32 = reg.Aproduct(b, precision=T.C64)
        assert O.rel_l2(y64, ref) < TOL64, nt
        assert O.rel_l2(y32, ref) < TOL32, nt
        reg.close()
        ys.append(y64)
    # a tile's transforms do not depend on the tile width: the same bits
    assert np.array_equal(ys[0], ys[1])
    monkeypatch.setenv("TDDA_XPASS", "0")
    lds = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    assert lds.xpass(T.C128) == "lds" and lds.xpass(T.C64) == "lds"
    z64 = lds.Aproduct(b)
    lds.close()
    # the two kernels round differently; both sit at FP64 accuracy
    assert O.rel_l2(y64, z64) < TOL64


def test_register_x_pass_solve(ctx, monkeypatch):
    # a BiCGSTAB solve through the register x pass reaches the true residual
    # and agrees with the solve through the LDS-staged pass
    Nx, Ny, Nz, d, lam = 40, 6, 4, 10.0, 500.0
    R = O.build_a_bulk(Nx, Ny, Nz)
    N = len(R) // 3
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    t = np.random.default_rng(11).random(N).repeat(3)
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
    E = O.einc(R, K, d, 1.0, nK, nE)
    P = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TDDA_XPASS", mode)
        core = T.AProductCore(ctx, R, Nx, Ny, Nz, d, lam, grid=(96, 12, 8))
        m = T.DeviceModel(core, 1)
        m.set(T.VEC_AL, al)
        m.set(T.VEC_E, E)
        info = m.solve(mode=T.MODE_MIXED, tol=1e-8, x0_zero=1, true_residual=1)[0]
        assert info["status"] == T.CONVERGED and info["true_residual"] <= 1e-8, info
        P[mode] = m.get(T.VEC_P)[0]
        m.close()
        core.close()
    assert O.rel_l2(P["1"], P["0"]) < 1e-7


YCASES = [  # (Nx, Ny, Nz), grid, occupied fraction
    ((4, 16, 3), (8, 32, 8), 1.0),
    ((3, 23, 2), (8, 45, 8), 0.7),
    ((3, 38, 2), (8, 75, 8), 1.0),
    ((4, 40, 3), (8, 96, 8), 1.0),
    ((5, 80, 3), (10, 160, 8), 0.6),
    ((3, 90, 2), (8, 180, 8), 1.0),
    ((3, 113, 2), (8, 225, 8), 1.0),
    ((4, 128, 2), (8, 256, 8), 0.5),
]


@pytest.mark.parametrize("dims,grid,frac", YCASES)
def test_register_y_pass(ctx, monkeypatch, dims, grid, frac):
    # k_pyr (forward P2 and inverse P4, one LDS exchange each) against the
    # numpy FFT convolution and against the LDS-staged y passes
    Nx, Ny, Nz = dims
    print(f"register y pass {dims} grid {grid} frac {frac}", flush=True)
    R, b = problem(Nx, Ny, Nz, frac, seed=grid[1])
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, 10.0, O.wavenumber(500.0), b, grid=grid)
    monkeypatch.setenv("TDDA_YPASS", "1")
    reg = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    assert reg.ypass(T.C128) == "register" and reg.ypass(T.C64) == "register"
    y64 = reg.Aproduct(b)
    assert O.rel_l2(y64, ref) < TOL64
    assert O.rel_l2(reg.Aproduct(b, precision=T.C64), ref) < TOL32
    reg.close()
    monkeypatch.setenv("TDDA_YPASS", "0")
    lds = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    assert lds.ypass(T.C128) == "lds"
    z64 = lds.Aproduct(b)
    lds.close()
    assert O.rel_l2(y64, z64) < TOL64


LONG = [  # complex64 register kernels with factors above 16: (dims, grid, axis)
    ((63, 3, 2), (125, 8, 8), "x"),
    ((81, 4, 2), (162, 8, 8), "x"),
    ((100, 3, 2), (200, 8, 8), "x"),
    ((108, 2, 2), (216, 8, 8), "x"),
    ((122, 2, 2), (243, 8, 8), "x"),
    ((162, 2, 2), (324, 8, 8), "x"),
    ((180, 2, 2), (360, 8, 8), "x"),
    ((200, 2, 2), (400, 8, 8), "x"),
    ((203, 2, 2), (405, 8, 8), "x"),
    ((243, 2, 2), (486, 8, 8), "x"),
    ((240, 2, 2), (480, 8, 8), "x"),
    ((250, 2, 2), (500, 8, 8), "x"),
    ((270, 2, 2), (540, 8, 8), "x"),
    ((2, 200, 2), (8, 400, 8), "y"),
    ((2, 81, 2), (8, 162, 8), "y"),
    ((2, 125, 2), (8, 250, 8), "y"),
    ((135, 3, 2), (270, 8, 8), "x"),
    ((256, 3, 2), (512, 8, 8), "x"),
    ((313, 2, 2), (625, 8, 8), "x"),
    ((365, 2, 2), (729, 8, 8), "x"),
    ((480, 2, 2), (960, 8, 8), "x"),
    ((512, 2, 2), (1024, 8, 8), "x"),
    ((486, 2, 2), (972, 8, 8), "x"),
    ((500, 2, 2), (1000, 8, 8), "x"),
    ((2, 500, 2), (8, 1000, 8), "y"),
    ((3, 135, 2), (8, 270, 8), "y"),
    ((2, 300, 2), (8, 600, 8), "y"),
    ((2, 512, 2), (8, 1024, 8), "y"),
]


# complex128 register kernels with a factor above 16: the lengths up to 540
# whose transforms fit the registers of two or three waves
MID128 = {125, 162, 200, 216, 243, 250, 270, 288, 300, 320, 324, 360, 375, 384, 400, 405, 432, 450, 480, 486,
          500, 512, 540}


@pytest.mark.parametrize("dims,grid,axis", LONG)
def test_register_passes_long_complex64(ctx, monkeypatch, dims, grid, axis):
    # lengths 125-1024: the complex64 x and y passes run k_pxr / k_pyr with
    # 512-thread workgroups and factors up to 40; complex128 runs them for
    # the lengths of MID128 and keeps the LDS-staged kernels for the others
    Nx, Ny, Nz = dims
    print(f"long register {axis} pass {dims} grid {grid}", flush=True)
    R, b = problem(Nx, Ny, Nz, 1.0, seed=grid[0] + grid[1])
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, 10.0, O.wavenumber(500.0), b, grid=grid)
    monkeypatch.setenv("TDDA_XPASS", "1")
    monkeypatch.setenv("TDDA_YPASS", "1")
    reg = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    kind = reg.xpass if axis == "x" else reg.ypass
    L = grid[0] if axis == "x" else grid[1]
    assert kind(T.C64) == "register" and kind(T.C128) == ("register" if L in MID128 else "lds")
    y32 = reg.Aproduct(b, precision=T.C64)
    assert O.rel_l2(y32, ref) < TOL32
    y64 = reg.Aproduct(b)
    assert O.rel_l2(y64, ref) < TOL64
    reg.close()
    monkeypatch.setenv("TDDA_XPASS", "0")
    monkeypatch.setenv("TDDA_YPASS", "0")
    lds = T.AProductCore(ctx, R, Nx, Ny, Nz, 10.0, 500.0, grid=grid)
    z32 = lds.Aproduct(b, precision=T.C64)
    z64 = lds.Aproduct(b)
    lds.close()
    assert O.rel_l2(y32, z32) < 2 * TOL32
    assert O.rel_l2(y64, z64) < TOL64
