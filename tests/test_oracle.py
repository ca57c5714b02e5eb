"""CPU tests that pin the oracle (oracle/) before it is trusted as the checker.

Pins: the reference's own Get_Alpha / nEPerpinXZ / build_a_bulk compiled from
its tools.cpp (fixtures in tests/golden/alpha_ref.npz), long-double
evaluations of the textbook dyadic Green function, closed-form invariants of
the interaction tensor, long-double direct sums, numpy's FFT and dense
linear solves.  A planted mistake in a copy of the tensor must fail the
invariant checks.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# --- polarizability, polarisation, lattice order (tools.cpp) ---------------
def test_get_alpha_matches_reference_tools_cpp():
    g = gold("alpha_ref.npz")
    for i in range(len(g["lam"])):
        a = O.get_alpha(g["K"][i], g["d"][i], complex(g["diel"][i]), g["nE"][i], g["nK"][i])
        assert abs(a - g["alpha"][i]) <= 1e-15 * abs(g["alpha"][i])


def test_get_alpha_live_reference_when_built():
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources not mounted)")
    out = np.zeros(2)
    nE = np.array([np.cos(0.3), 0, -np.sin(0.3)])
    nK = np.array([np.sin(0.3), 0, np.cos(0.3)])
    for diel in (1.01, 2.25 + 0.01j, 3.9999 + 0.04j):
        ref.ref_get_alpha(500.0, 2 * np.pi / 500, 10.0, complex(diel).real, complex(diel).imag,
                          O._ptr(nE), O._ptr(nK), O._ptr(out))
        a = O.get_alpha(2 * np.pi / 500, 10.0, complex(diel), nE, nK)
        assert abs(a - complex(out[0], out[1])) <= 1e-15 * abs(a)


def test_get_alpha_closed_form_limits():
    # eps = 1 gives alpha = 0 (the reason for the +0.01 rule, AProductCore.cpp:19-24)
    assert O.get_alpha(0.01, 10.0, 1.0 + 0j, [1, 0, 0], [0, 0, 1]) == 0
    # static limit K -> 0: Clausius-Mossotti a1 = 3 d^3 /(4 pi) (eps-1)/(eps+2)
    d, eps = 10.0, 2.25
    a = O.get_alpha(1e-9, d, eps, [1, 0, 0], [0, 0, 1])
    cm = 3 * d ** 3 / (4 * np.pi) * (eps - 1) / (eps + 2)
    assert abs(a - cm) < 1e-12 * cm


def test_nE_perp_and_bulk_match_reference():
    g = gold("alpha_ref.npz")
    for th, v in zip(g["thetas"], g["nperp"]):
        assert np.allclose(O.ne_perp_in_xz(th), v, atol=1e-15)
    assert np.array_equal(O.build_a_bulk(2, 3, 4), g["bulk234"])


def test_material_rule_and_wavenumber():
    m = O.material_rule(np.array([1.0, 2.25 + 0.01j, 0.5]), nback=1.0)
    assert m[0] == 1.01 and m[1] == 2.25 + 0.01j and m[2] == 0.51
    m2 = O.material_rule(np.array([2.25 * 1.44]), nback=1.2)
    assert abs(m2[0] - 2.25) < 1e-15
    assert O.wavenumber(500.0, 1.0) == 2 * np.pi / 500.0
    assert abs(O.wavenumber(500.0, 1.5) - 2 * np.pi * 1.5 / 500.0) < 1e-18


def test_diel_interpolation_and_local_field():
    mat = np.array([1.01 + 0j, 2.25 + 0.01j])
    t = np.array([0.0, 1.0, 0.5, 0.25])
    diel, al = O.diel_al(t, mat, 500.0, 2 * np.pi / 500, 10.0, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]))
    assert diel[0] == mat[0] and diel[1] == mat[1]
    assert abs(diel[2] - (mat[0] + mat[1]) / 2) < 1e-15
    assert abs(diel[3] - (0.75 * mat[0] + 0.25 * mat[1])) < 1e-15
    for i in range(4):
        a = O.get_alpha(2 * np.pi / 500, 10.0, complex(diel[i]), [1.0, 0, 0], [0, 0, 1.0])
        assert abs(al[i] - 1 / a) < 1e-14 * abs(al[i])
    P = np.array([1 + 2j, -1j, 0.5, 3.0])
    ei = O.local_field(al, P, diel)
    assert np.allclose(ei, al * P * 3 / (2 + diel), rtol=1e-15)


def test_einc_closed_form():
    R = O.build_a_bulk(3, 2, 4)
    K, d = 2 * np.pi / 500, 10.0
    th = np.deg2rad(20)
    nK = np.array([np.sin(th), 0, np.cos(th)])
    nE = O.ne_perp_in_xz(th)
    E = O.einc(R, K, d, 2.0, nK, nE).reshape(-1, 3)
    Rr = R.reshape(-1, 3)
    ph = K * d * (Rr @ nK)
    assert np.allclose(E, 2.0 * nE[None, :] * np.exp(1j * ph)[:, None], rtol=0, atol=1e-15)


# --- interaction tensor (AProductCore.cpp:588-623) --------------------------
def _green_invariant_errors(green_fn, K, pts):
    """Largest violation of the closed-form invariants of A = -G."""
    worst = 0.0
    for (x, y, z) in pts:
        A = green_fn(K, x, y, z)
        M = np.array([[A[0], A[1], A[2]], [A[1], A[3], A[4]], [A[2], A[4], A[5]]])
        r = np.sqrt(x * x + y * y + z * z)
        u = np.array([x, y, z]) / r
        ph = np.exp(1j * K * r)
        scale = abs(K * K * ph / r) + abs(ph / r ** 3)
        worst = max(worst, abs(np.trace(M) - (-2 * K * K * ph / r)) / scale)
        worst = max(worst, np.max(np.abs(M @ u - 2 * (1j * K * r - 1) * ph / r ** 3 * u)) / scale)
        # transverse part: (I - uu) A (I - uu) = -k^2 e^{ikr}/r (I - uu) - (ikr-1)e^{ikr}/r^3 (I - uu)
        Pt = np.eye(3) - np.outer(u, u)
        T = Pt @ M @ Pt
        Tref = -(K * K * ph / r + (1j * K * r - 1) * ph / r ** 3) * Pt
        worst = max(worst, np.max(np.abs(T - Tref)) / scale)
        Am = green_fn(K, -x, -y, -z)
        worst = max(worst, np.max(np.abs(Am - A)) / scale)
    return worst


PTS = [(10.0, 0, 0), (0, 10.0, 0), (0, 0, 10.0), (10.0, 20.0, -30.0), (-7.0, 3.0, 5.5), (150.0, -80.0, 20.0)]


def test_green_matches_long_double_fixture():
    g = gold("green_ld.npz")
    for off, A in zip(g["offs"], g["A"]):
        a = O.ldr_inter(float(g["K"]), *off)
        assert np.max(np.abs(a - A)) <= 1e-14 * np.max(np.abs(A))
    assert np.all(O.ldr_inter(0.01, 0.0, 0.0, 0.0) == 0)


def test_green_closed_form_invariants():
    for K in (2 * np.pi / 500, 2 * np.pi / 450 * 1.5):
        assert _green_invariant_errors(O.ldr_inter, K, PTS) < 1e-14


def test_planted_green_mistakes_are_caught():
    """Mistakes planted in a copy of the tensor must fail the invariants."""
    def swapped(K, x, y, z):      # transposed operand: xz <-> yz
        a = O.ldr_inter(K, x, y, z)
        return a[[0, 1, 4, 3, 2, 5]]

    def sign(K, x, y, z):         # wrong sign of the off-diagonal terms
        a = O.ldr_inter(K, x, y, z)
        a[[1, 2, 4]] *= -1
        return a

    def dropped(K, x, y, z):      # dropped radiative k^2 term
        r2 = x * x + y * y + z * z
        r = np.sqrt(r2)
        c1 = np.exp(1j * K * r) / r ** 3
        c2 = (1 - 1j * K * r) / r2
        c3 = -3 * c2
        return c1 * np.array([c2 * (r2 - 3 * x * x), x * y * c3, z * x * c3, c2 * (r2 - 3 * y * y),
                              y * z * c3, c2 * (r2 - 3 * z * z)])

    K = 2 * np.pi / 500
    for bad in (swapped, sign, dropped):
        assert _green_invariant_errors(bad, K, PTS) > 1e-6


# --- FFT and operator --------------------------------------------------------
# 101, 127 and 2 * 67: the reference's 2N - 1 grids have prime factors above 64
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 8, 12, 30, 31, 64, 160, 243, 101, 127, 134])
def test_fft1d_matches_numpy(L):
    x = np.random.default_rng(L).standard_normal(L) + 1j * np.random.default_rng(L + 1).standard_normal(L)
    assert np.allclose(O.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * max(1, np.abs(x).sum()))
    assert np.allclose(O.fft1d(x, +1), np.fft.ifft(x) * L, rtol=0, atol=1e-13 * max(1, np.abs(x).sum()))


def test_aproduct_direct_and_fft_match_long_double_fixture():
    g = gold("aproduct_tiny.npz")
    for name in ("a", "b", "c"):
        Nx, Ny, Nz = (int(v) for v in g[f"{name}_dims"])
        R, d, K, b, y = (g[f"{name}_{k}"] for k in ("R", "d", "K", "b", "y"))
        d, K = float(d), float(K)
        assert O.rel_l2(O.aproduct_direct(R, d, K, b), y) < 1e-14
        for grid in (None, (2 * Nx, 2 * Ny, 2 * Nz), (2 * Nx + 3, 2 * Ny + 1, 2 * Nz + 2)):
            op = O.OrcOperator(R, Nx, Ny, Nz, d, K, grid=grid)
            assert O.rel_l2(op.apply(b), y) < 1e-14
        assert O.rel_l2(O.aproduct_numpy_fft(R, Nx, Ny, Nz, d, K, b), y) < 1e-14


def test_aproduct_is_complex_symmetric():
    # A^T = A (not Hermitian): y^T (A x) = x^T (A y) -- the adjoint solve reuses A
    rng = np.random.default_rng(3)
    R = O.build_a_bulk(5, 4, 3)
    N = len(R) // 3
    op = O.OrcOperator(R, 5, 4, 3, 10.0, 2 * np.pi / 500)
    x = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    y = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N)
    lhs = y @ op.apply(x)
    rhs = x @ op.apply(y)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


def test_dense_assembly_matches_operator():
    R = O.build_a_bulk(3, 3, 2)
    N = len(R) // 3
    K = 2 * np.pi / 542
    A = O.assemble_dense(R, 10.0, K)
    b = np.random.default_rng(2).standard_normal(3 * N) + 0.5j
    assert O.rel_l2(A @ b, O.aproduct_direct(R, 10.0, K, b)) < 1e-14


# --- BiCGSTAB (DDAModel.cpp:131-260) ------------------------------------------
def test_bicgstab_matches_dense_solve():
    g = gold("solve_tiny.npz")
    for name in ("a", "b", "c"):
        Nx, Ny, Nz = (int(v) for v in g[f"{name}_dims"])
        R = O.build_a_bulk(Nx, Ny, Nz)
        K = O.wavenumber(float(g[f"{name}_lam"]))
        op = O.OrcOperator(R, Nx, Ny, Nz, float(g[f"{name}_d"]), K)
        P, st, it, err = op.bicgstab(g[f"{name}_al"], g[f"{name}_E"], tol=1e-13, max_it=2000)
        assert st == 0 and err <= 1e-13
        assert O.rel_l2(P, g[f"{name}_P"]) < 1e-11


def test_bicgstab_breakdown_and_maxit():
    R = O.build_a_bulk(3, 3, 3)
    N = len(R) // 3
    K = O.wavenumber(500.0)
    op = O.OrcOperator(R, 3, 3, 3, 10.0, K)
    mat = O.material_rule(np.array([1.0, 2.25]))
    _, al = O.diel_al(np.ones(3 * N), mat, 500.0, K, 10.0, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]))
    E = O.einc(R, K, 10.0, 1.0, np.array([0, 0, 1.0]), np.array([1.0, 0, 0]))
    # zero right-hand side: r0.rl == 0 at the first loop test (DDAModel.cpp:193)
    P, st, it, err = op.bicgstab(al, np.zeros_like(E))
    assert st == 1 and it == 1
    # iteration cap: "does not converge" (DDAModel.cpp:256-259)
    P, st, it, err = op.bicgstab(al, E, max_it=1, tol=1e-30)
    assert st == 2 and it == 1
