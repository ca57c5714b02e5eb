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
