"""CPU tests of the C ABI: the HIP library loads, exports every symbol that
include/topodda.h declares, and fails cleanly (an error code, no crash) when
no GPU is present."""
import os
import re

import numpy as np

import topodda as T
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(tdda_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = T.lib()
    decl = declared_symbols()
    assert len(decl) >= 15
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert set(T.EXPORTS) <= set(decl)


def test_host_green_tensor_matches_oracle():
    # tdda_green_tensor is host code (AProductCore::LDR_inter replacement)
    K = 2 * np.pi / 500
    for off in [(10.0, 0, 0), (10.0, 20.0, -30.0), (-7.0, 3.0, 5.5), (0.0, 0.0, 0.0)]:
        a = T.green_tensor(K, *off)
        b = O.ldr_inter(K, *off)
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1e-300, np.max(np.abs(b)))


def test_errors_are_codes_not_crashes():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    try:
        T.Context(0)
    except T.TddaError as e:
        assert "error" in str(e)
    else:
        raise AssertionError("context creation without a GPU must fail")


def test_host_setup_math_matches_oracle():
    # tdda_material_rule / get_alpha / incident_field / polarizability / local_field
    # are host code (DDAModel set-up); compared with the pinned oracle
    g = np.load(os.path.join(ROOT, "tests", "golden", "alpha_ref.npz"), allow_pickle=False)
    for i in range(len(g["lam"])):
        a = T.get_alpha(g["lam"][i], g["K"][i], g["d"][i], complex(g["diel"][i]), g["nE"][i], g["nK"][i])
        assert abs(a - g["alpha"][i]) <= 1e-15 * abs(g["alpha"][i])
    mat = np.array([1.0, 2.25 + 0.01j, 0.7])
    assert np.array_equal(T.material_rule(mat, 1.1), O.material_rule(mat, 1.1))
    R = T.build_a_bulk(4, 3, 5)
    assert np.array_equal(R, O.build_a_bulk(4, 3, 5))
    th = np.deg2rad(30)
    nK = np.array([np.sin(th), 0, np.cos(th)])
    nE = T.ne_perp_in_xz(th)
    assert np.allclose(nE, g["nperp"][list(np.round(np.rad2deg(g["thetas"]))).index(30.0)], atol=1e-15)
    K = T.wavenumber(542.0)
    assert np.array_equal(T.incident_field(R, K, 10.0, nK, nE), O.einc(R, K, 10.0, 1.0, nK, nE))
    t = np.random.default_rng(0).random(len(R))
    m = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    d1, a1 = T.polarizability(t, m, 542.0, K, 10.0, nE, nK)
    d2, a2 = O.diel_al(t, m, 542.0, K, 10.0, nE, nK)
    assert np.allclose(d1, d2, rtol=1e-15, atol=0) and np.allclose(a1, a2, rtol=1e-14, atol=0)
    P = np.random.default_rng(1).standard_normal(len(R)) + 1j
    assert np.allclose(T.local_field(a1, P, d1), O.local_field(a2, P, d2), rtol=1e-14, atol=0)


def test_material_rule_threshold_of_each_constructor():
    # AProductCore.cpp:19-24 adds 0.01 when |eps| <= 1.01; the periodic
    # constructor (AProductCore.cpp:261-267) only when |eps| < 1.01, so a
    # material at exactly 1.01 differs between the two
    mat = np.array([1.01, 1.0, 1.0100000000000002, 0.5 + 0.8j, 0.5 + 0.9j, 2.25 + 0.01j])
    plain = T.material_rule(mat)
    per = T.material_rule(mat, periodic=True)
    assert plain[0] == 1.02 and per[0] == 1.01
    assert plain[1] == per[1] == 1.01
    assert plain[2] == per[2] == mat[2]              # just above 1.01: kept
    assert plain[3] == per[3] == mat[3] + 0.01       # |eps| = 0.943
    assert plain[4] == per[4] == mat[4]              # |eps| = 1.0296
    assert plain[5] == per[5] == mat[5]
    for periodic in (False, True):
        for nback in (1.0, 1.3):
            assert np.array_equal(T.material_rule(mat, nback, periodic), O.material_rule(mat, nback, periodic))


def test_op_plan_grid_lengths_are_valid():
    # every planner length is 2,3,5-smooth, at least 2N - 1 (the linear
    # convolution) and at most 30 % above it (the x-pass cost table's window)
    def smooth(n):
        for p in (2, 3, 5):
            while n % p == 0:
                n //= p
        return n == 1
    for n in list(range(2, 140)) + list(range(140, 700, 7)):
        g = T.op_plan(n, n if n < 300 else 8, 3)
        for L, N in zip(g, (n, n if n < 300 else 8, 3)):
            assert smooth(L) and L >= 2 * N - 1, (n, g)
        assert g[0] <= 1.3 * (2 * n - 1) + 1, (n, g)


def test_op_plan_grid_and_32bit_guard():
    # grids the planner picks: y and z the smallest 2,3,5-smooth >= 2N - 1
    # (y: 75 and 125 skipped), x among the lengths within 30 % above 2Nx - 1
    # the one whose slower precision is closest to its best x-pass time, if
    # more than 5 % better than the smallest smooth length
    assert T.op_plan(16, 16, 16) == (32, 32, 32)
    assert T.op_plan(63, 63, 12) == (128, 128, 24)  # 125 = 5^3: 1.5x / 2.4x the x-pass time of 128
    assert T.op_plan(38, 38, 8)[1:] == (80, 15)
    assert T.op_plan(48, 20, 5) == (100, 40, 9)     # 96 = 12 x 8 costs more than 100 = 10 x 10
    assert T.op_plan(100, 10, 4)[0] == 200
    assert T.op_plan(121, 10, 4)[0] == 256          # 243 = 27 x 9 and 250 = 25 x 10 cost more
    assert T.op_plan(200, 10, 4)[0] == 400
    assert T.op_plan(400, 10, 4)[0] == 800          # beyond the table: next smooth length
    assert T.op_plan(80, 80, 16) == (160, 160, 32)  # 162 measures within 5 % of 160: the smaller stays
    assert T.op_plan(128, 128, 32) == (256, 256, 64)
    assert T.op_plan(512, 512, 64) == (1024, 1024, 128)
    # the pass kernels use 32-bit offsets within one right-hand side's
    # Nx x My x Mz x 3 intermediate: larger structures are refused
    T.op_plan(40, 1000, 16, grid=(80, 2048, 4096))  # 40 * 2048 * 4096 * 3 = 1.01e9 < 2^31
    for args, grid in (((600, 600, 600), None), ((90, 1000, 16), (180, 2048, 4096))):  # 2.6e9, 2.26e9
        try:
            T.op_plan(*args, grid=grid)
        except T.TddaError as e:
            assert "2^31" in str(e)
        else:
            raise AssertionError("Nx * My * Mz * 3 >= 2^31 must be refused")
