"""GPU tests of the tile autotuner (tdda_op_autotune): the tuned tiles stay
within the bounds of the kernels it chose, the tuned operator still matches the oracle
(AProductCore.cpp:697-758) to the operator bar, and a tuned model gives the
same bits alone and in a batch with an untuned one."""
import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(24, 20, 8), (80, 30, 6)])
def test_autotuned_operator_matches_oracle(ctx, dims):
    Nx, Ny, Nz = dims
    d, lam = 10.0, 500.0
    print(f"autotune {dims}", flush=True)
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
    before = {p: (core.tiles(p), core.zy_fused(p)) for p in (T.C128, T.C64)}
    core.autotune(reps=10)
    Mz = core.grid[2]
    for p in (T.C128, T.C64):
        pz, pz1, tfy, tfx = core.tiles(p)
        (bz, bz1, by, bx), bf = before[p]
        # fused z + y passes: one z tile per x slab (Ny pencils); separate
        # passes: z tiles only narrow from the planner's; the x and y tiles
        # are those of the kernel the tuner chose (register-resident with 256
        # or 512 threads, or LDS-staged), each within one ky / x row of 3 Mz
        # columns
        if core.zy_fused(p):
            assert pz == pz1 == Ny
        elif not bf:
            assert 1 <= pz <= bz and 1 <= pz1 <= bz1
        else:
            assert pz >= 1 and pz1 >= 1
        assert 1 <= tfy <= 3 * Mz and 3 <= tfx <= 3 * Mz and tfx % 3 == 0
        assert core.xpass(p) in ("register", "lds") and core.ypass(p) in ("register", "lds")
    R = O.build_a_bulk(Nx, Ny, Nz)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(len(R)) + 1j * rng.standard_normal(len(R))
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, d, O.wavenumber(lam), b, grid=core.grid)
    assert O.rel_l2(core.Aproduct(b), ref) < 1e-12
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < 5e-6
    core.close()


def test_tuned_model_alone_equals_batched(ctx):
    specs = [(24, 20, 8, 10.0, 500.0), (12, 7, 5, 10.0, 542.0)]
    cores, models, data = [], [], []
    for i, (Nx, Ny, Nz, d, lam) in enumerate(specs):
        core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
        if i == 0:
            core.autotune(reps=10)
        R = O.build_a_bulk(Nx, Ny, Nz)
        K = O.wavenumber(lam)
        t = np.random.default_rng(i).random(len(R) // 3).repeat(3)
        nK, nE = np.array([0, 0, 1.0]), np.array([1.0, 0, 0])
        _, al = O.diel_al(t, O.material_rule(np.array([1.0, 2.25 + 0.01j])), lam, K, d, nE, nK)
        E = O.einc(R, K, d, 1.0, nK, nE)
        m = T.DeviceModel(core, 1)
        m.set(T.VEC_AL, al)
        m.set(T.VEC_E, E)
        m.set(T.VEC_P, np.zeros_like(E))
        cores.append(core)
        models.append(m)
        data.append((al, E))
    b = T.Batch(models)
    infos = b.solve(mode=T.MODE_MIXED, tol=1e-8)
    for core, m, (al, E), ib in zip(cores, models, data, infos):
        one = T.DeviceModel(core, 1)
        one.set(T.VEC_AL, al)
        one.set(T.VEC_E, E)
        one.set(T.VEC_P, np.zeros_like(E))
        i1 = one.solve(mode=T.MODE_MIXED, tol=1e-8)[0]
        assert i1["status"] == T.CONVERGED and i1["iterations"] == ib["iterations"]
        assert np.array_equal(one.get(T.VEC_P)[0], m.get(T.VEC_P)[0])
        one.close()
    b.close()
    for m in models:
        m.close()
    for c in cores:
        c.close()


def test_tuner_keeps_its_y_pass_decision(ctx, monkeypatch):
    # TDDA_TUNE_YLDS=1 makes the LDS-staged y kernels win the y step; the
    # x step that follows must not drop that decision.  The y step runs
    # for separate z and y passes (TDDA_TUNE_NOZY=1 makes them win over the
    # fused ones, which keep the register four-step)
    Nx, Ny, Nz, d, lam = 24, 20, 8, 10.0, 500.0
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
    assert core.ypass(T.C128) == "register" and core.ypass(T.C64) == "register"
    monkeypatch.setenv("TDDA_TUNE_YLDS", "1")
    monkeypatch.setenv("TDDA_TUNE_NOZY", "1")
    core.autotune(reps=5)
    monkeypatch.delenv("TDDA_TUNE_YLDS")
    monkeypatch.delenv("TDDA_TUNE_NOZY")
    assert not core.zy_fused(T.C128) and not core.zy_fused(T.C64)
    assert core.ypass(T.C128) == "lds" and core.ypass(T.C64) == "lds"
    R = O.build_a_bulk(Nx, Ny, Nz)
    b = np.random.default_rng(5).standard_normal(len(R)) + 0.5j
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, d, O.wavenumber(lam), b, grid=core.grid)
    assert O.rel_l2(core.Aproduct(b), ref) < 1e-12
    core.close()


def test_autotune_for_several_right_hand_sides(ctx):
    # tdda_op_autotune_ex: the tuner's scratch model carries nrhs right-hand
    # sides; the tuned operator still matches the oracle for every one
    Nx, Ny, Nz, d, lam = 24, 20, 10, 10.0, 500.0
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam)
    core.autotune(reps=5, nrhs=4)
    R = O.build_a_bulk(Nx, Ny, Nz)
    rng = np.random.default_rng(8)
    b = rng.standard_normal((4, len(R))) + 1j * rng.standard_normal((4, len(R)))
    y = core.Aproduct(b)
    for j in range(4):
        ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, d, O.wavenumber(lam), b[j], grid=core.grid)
        assert O.rel_l2(y[j], ref) < 1e-12
    for bad in (0, 5000):
        with pytest.raises(T.TddaError):
            core.autotune(reps=5, nrhs=bad)
    core.close()


def test_autotune_keeps_wide_complex128_register_x_pass(ctx):
    # x length 162 = 18 x 9: a factor above 16, which the register kernels
    # run with 512 threads in both precisions; the tuner must time that
    # kernel for complex128 too (the LDS-staged one is twice as slow)
    Nx, Ny, Nz, d, lam = 81, 32, 12, 10.0, 500.0
    core = T.AProductCore(ctx, None, Nx, Ny, Nz, d, lam, grid=(162, 0, 0))
    core.autotune(reps=5)
    assert core.xpass(T.C128) == "register" and core.xpass(T.C64) == "register"
    R = O.build_a_bulk(Nx, Ny, Nz)
    b = np.random.default_rng(9).standard_normal(len(R)) + 0.25j
    ref = O.aproduct_numpy_fft(R, Nx, Ny, Nz, d, O.wavenumber(lam), b, grid=core.grid)
    assert O.rel_l2(core.Aproduct(b), ref) < 1e-12
    assert O.rel_l2(core.Aproduct(b, precision=T.C64), ref) < 5e-6
    core.close()
