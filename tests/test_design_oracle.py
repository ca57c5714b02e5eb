"""CPU tests that pin the oracle's restatement of the design update of an
extruded structure: CoreStructure::UpdateStr (CoreStructure.cpp:31-104) with
SpacePara's cone filter (SpacePara.cpp:37-48, 1241-1270), SmoothDensity
(FilterOption.cpp:49-60) and the beta schedules (tools.cpp:776-798).

Pins: fixtures written by the reference's own classes compiled from its
sources (tests/golden/design_ref.npz, tests/golden/make_golden.py), the live
reference when oracle/_ref is built, and properties the filter must have
(a constant design stays constant, a radius below one lattice step is the
identity, beta = 0 projects nothing, the projection fixes 0, eta and 1).
A planted mistake (an unnormalised filter) must fail them.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    g = np.load(os.path.join(GOLD, "design_ref.npz"), allow_pickle=False)
    k = 0
    while f"c{k}_case" in g:
        yield g, k
        k += 1


def test_design_update_matches_reference_fixture():
    n = 0
    for g, k in cases():
        Nx, Ny, Nz, r, beta, eta = g[f"c{k}_case"]
        Nx, Ny, Nz = int(Nx), int(Ny), int(Nz)
        o, f, p = O.design_update(g[f"c{k}_p0"], g[f"c{k}_step"], Nx, Ny, r, beta, eta)
        assert np.array_equal(f, g[f"c{k}_filtered"])
        assert np.max(np.abs(p - g[f"c{k}_para"])) <= 1e-15
        assert np.array_equal(O.design_diel_old(p, Nx, Ny, Nz), O.design_diel_old(g[f"c{k}_para"], Nx, Ny, Nz))
        assert np.max(np.abs(O.design_diel_old(p, Nx, Ny, Nz) - g[f"c{k}_diel"])) <= 1e-15
        n += 1
    assert n == 4


def test_beta_schedules_match_reference_fixture():
    g = np.load(os.path.join(GOLD, "design_ref.npz"), allow_pickle=False)
    for kind in range(3):
        got = [O.beta_update(kind, x, 10.0, 1.0, 200.0) for x in g["beta_x"]]
        assert np.array_equal(np.array(got), g["beta"][kind])


def test_design_update_live_reference_when_built():
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources not mounted)")
    rng = np.random.default_rng(5)
    for (Nx, Ny, Nz, r, beta, eta) in [(8, 5, 2, 2.5, 4.0, 0.5), (4, 4, 1, 1.5, 0.0, 0.4)]:
        p0 = rng.random(Nx * Ny)
        st = 0.3 * rng.standard_normal(Nx * Ny)
        diel, f, p = np.zeros(3 * Nx * Ny * Nz), np.zeros(Nx * Ny), np.zeros(Nx * Ny)
        assert ref.ref_design_update(Nx, Ny, Nz, O._ptr(p0), O._ptr(st), r, beta, eta, O._ptr(diel), O._ptr(f),
                                     O._ptr(p)) == 0
        o, f2, p2 = O.design_update(p0, st, Nx, Ny, r, beta, eta)
        assert np.array_equal(f2, f)
        assert np.max(np.abs(p2 - p)) <= 1e-15


def test_cone_filter_properties():
    Nx, Ny = 7, 5
    # a constant design is a fixed point of the normalised filter
    c = np.full(Nx * Ny, 0.37)
    _, f, p = O.design_update(c, None, Nx, Ny, 2.7, 0.0, 0.5)
    assert np.max(np.abs(f - 0.37)) < 1e-15 and np.array_equal(p, f)
    # r^2 - 0.01 < 1: each column sees only itself
    rng = np.random.default_rng(1)
    x = rng.random(Nx * Ny)
    assert np.array_equal(O.design_update(x, None, Nx, Ny, 1.0, 0.0, 0.5)[1], x)
    # r = 2: the 3 x 3 neighbourhood with weights 2, 1, 2 - sqrt(2) (interior column)
    w = O.cone_weights(Nx, Ny, 2.0)[3 * Ny + 2]
    got = {j: round(v, 12) for j, v in w}
    want = {}
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            want[(3 + dx) * Ny + 2 + dy] = round(2.0 - np.hypot(dx, dy), 12)
    assert got == want
    # steps are clamped to [0, 1] before filtering
    o, _, _ = O.design_update(np.full(4, 0.5), np.array([0.7, -0.9, 0.2, 0.0]), 2, 2, 1.0, 0.0, 0.5)
    assert np.array_equal(o, [1.0, 0.0, 0.7, 0.5])


def test_smooth_density_properties():
    for beta in (0.0, 1.0, 20.0):
        for eta in (0.3, 0.5, 0.7):
            assert abs(O.smooth_density(0.0, beta, eta)) < 1e-15
            assert abs(O.smooth_density(1.0, beta, eta) - 1.0) < 1e-15
            assert abs(O.smooth_density(eta, beta, eta) - eta) < 1e-15
    xs = np.linspace(0, 1, 11)
    assert np.allclose([O.smooth_density(x, 0.0, 0.5) for x in xs], xs, atol=1e-15)
    # steep projection pushes values away from eta
    assert O.smooth_density(0.4, 50.0, 0.5) < 0.01 and O.smooth_density(0.6, 50.0, 0.5) > 0.99
    with pytest.raises(ValueError):
        O.smooth_density(1.2, 1.0, 0.5)


def test_planted_unnormalised_filter_is_caught():
    # a filter that forgets the division by the weights' sum fails the
    # constant-design property
    def bad(x, Nx, Ny, r):
        return np.array([sum(w * x[j] for j, w in row) for row in O.cone_weights(Nx, Ny, r)])
    c = np.full(12, 0.5)
    assert np.max(np.abs(bad(c, 4, 3, 2.0) - 0.5)) > 0.1
