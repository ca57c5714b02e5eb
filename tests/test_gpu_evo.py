"""GPU test of the reference-signature optimiser (EvoDDAModel +
EvoOptimization, EvoDDAModel.cpp:393-765) over the host classes, driven as
the reference's "EvoOpt 2D input" task drives it (main.cpp:291-371): three
design iterations on an extruded 8 x 8 x 3 structure (column parameters,
cone filter r = 2, exp beta schedule, PointE objective, Adam), solves on the
device.  The same loop is restated here on the oracle -- its BiCGSTAB
(DDAModel.cpp:131-260), PointE value and closed-form adjoint source (pinned
by differentiation), the filtered design update (pinned bitwise against the
reference's CoreStructure::UpdateStr) -- and the objective history and final
design values must agree."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "topo-dda-forwin64_amd", "dda_evo_check")


def pto_f(x, beta, eta):
    """d SmoothDensity / dx (EvoDDAModel.cpp:1163-1175)."""
    if 0.0 <= x <= eta:
        return beta * math.exp(-beta * (1 - x / eta)) + math.exp(-beta)
    return beta * math.exp(-beta * (x - eta) / (1 - eta)) + math.exp(-beta)


def test_pto_f_is_the_derivative_of_smooth_density():
    # central differences of the oracle's SmoothDensity away from eta
    for beta in (1.0, 8.0, 30.0):
        for x in (0.05, 0.3, 0.45, 0.55, 0.8, 0.97):
            h = 1e-6
            fd = (O.smooth_density(x + h, beta, 0.5) - O.smooth_density(x - h, beta, 0.5)) / (2 * h)
            assert abs(fd - pto_f(x, beta, 0.5)) <= 1e-6 * max(1.0, abs(fd))


def evo_restated(Nx, Ny, Nz, d, lam, col0, iters, tol):
    R = O.build_a_bulk(Nx, Ny, Nz)
    K = O.wavenumber(lam)
    mat = O.material_rule(np.array([1.0, 2.25 + 0.01j]))
    nK, nE = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    r0 = np.array([0.5 * d * (Nx - 1), 0.5 * d * (Ny - 1), d * (Nz + 4)])
    op = O.OrcOperator(R, Nx, Ny, Nz, d, K)
    sites = R.reshape(-1, 3)
    col_of = sites[:, 0] * Ny + sites[:, 1]
    ncol = Nx * Ny
    origin, filtered, para = col0.copy(), col0.copy(), col0.copy()
    rf, bmin, bmax, eta = 2.0, 1.0, 50.0, 0.5
    W = O.cone_weights(Nx, Ny, rf)
    b1 = 0.9
    b2 = 1 - (1 - b1) * (1 - b1)
    V = S = None
    hist = []
    for it in range(iters):
        t = np.repeat(para[col_of], 3)
        _, al = O.diel_al(t, mat, lam, K, d, nE, nK)
        E = O.einc(R, K, d, 1.0, nK, nE)
        P, st, _, _ = op.bicgstab(al, E, max_it=100000, tol=tol)
        assert st == 0
        hist.append(O.obj_point_e(R, d, K, P, r0, nK, nE))
        # devx = 0 for PointE; dA/dx P from the polarizability at t +- 0.001
        eps = 0.001
        sign = np.where(para >= eps, -1.0, 1.0)
        _, al1 = O.diel_al(np.repeat((para + sign * eps)[col_of], 3), mat, lam, K, d, nE, nK)
        Adevxp = (al1 - al) / np.repeat(sign[col_of] * eps, 3) * P
        src = O.devp_point_e(R, d, K, P, r0, nK, nE)
        lamv, st, _, _ = op.bicgstab(al, src, max_it=100000, tol=tol)
        assert st == 0
        grad = np.zeros(ncol)
        prod = (lamv * Adevxp).reshape(-1, 3).sum(axis=1)
        np.add.at(grad, col_of, -prod.real)
        if it >= 1:
            beta = O.beta_update(0, it - 1, iters - 1, bmin, bmax)
            res = np.zeros(ncol)
            for row in W:
                sw = sum(w for _, w in row)
                for j, w in row:
                    res[j] += w / sw
            grad = res * np.array([pto_f(v, beta, eta) for v in filtered]) * grad
        c1, c2 = 1 - b1 ** (it + 1), 1 - b2 ** (it + 1)
        V = (1 - b1) * grad / c1 if V is None else b1 * V + (1 - b1) * grad / c1
        S = (1 - b2) * grad ** 2 / c2 if S is None else b2 * S + (1 - b2) * grad ** 2 / c2
        grad = V / (np.sqrt(S) + 1e-8)
        eps_final = 0.1 if it <= 3 else 1.0
        mean = np.mean(np.abs(grad))
        if mean < 0.1:
            grad = grad / (mean / 0.1)
        beta = O.beta_update(0, it, iters - 1, bmin, bmax)
        origin, filtered, para = O.design_update(origin, eps_final * grad, Nx, Ny, rf, beta, eta)
    op.close()
    return np.array(hist), origin, filtered, para


def test_evo_optimization_matches_restated_loop(tmp_path):
    Nx, Ny, Nz, d, lam, seed, iters, tol = 8, 8, 3, 10.0, 500.0, 11, 3, 1e-12
    print("dda_evo_check 8x8x3, 3 iterations", flush=True)
    p = subprocess.run([EXE, str(tmp_path), str(Nx), str(Ny), str(Nz), str(d), str(lam), str(seed), str(iters),
                        str(tol)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    v = np.array([float(x) for x in (tmp_path / "evo.txt").read_text().split()])
    ncol = Nx * Ny
    col0 = v[:ncol]
    hist = v[ncol:ncol + iters]
    origin, filtered, para = (v[ncol + iters + k * ncol: ncol + iters + (k + 1) * ncol] for k in range(3))
    h2, o2, f2, p2 = evo_restated(Nx, Ny, Nz, d, lam, col0, iters, tol)
    assert np.all(np.abs(hist - h2) <= 1e-8 * np.abs(h2)), (hist, h2)
    assert np.max(np.abs(origin - o2)) <= 1e-7
    assert np.max(np.abs(filtered - f2)) <= 1e-7
    assert np.max(np.abs(para - p2)) <= 1e-6
    # the design moved, and the reference's result files are there
    assert np.max(np.abs(origin - col0)) > 0.05
    for f in ("convergence.txt", "Originiterations.txt", "Adjointiterations.txt",
              "CoreStructure/CoreStructure0.txt", "Model_output/Model_results0it2.txt"):
        assert (tmp_path / f).exists(), f
    conv = (tmp_path / "convergence.txt").read_text().split()
    assert len(conv) == iters and abs(float(conv[0]) - hist[0]) <= 1e-5 * hist[0]
