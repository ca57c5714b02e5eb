"""Seeded synthetic DDA problems (product code, no oracle).

Geometry and materials follow the reference's data contract
(CoreStructure.h:12-50: R, Nx/Ny/Nz, d, diel_old per component) and its set-up
rules: materials through the AProductCore.cpp:19-24 rule (so a vacuum label
becomes eps = 1.01), full boxes in build_a_bulk order (tools.cpp:497-510),
2D cone filter with the reference's weights (SpacePara.cpp:37-48 calweight /
circlerange, normalised as CoreStructure.cpp:64-77), extruded along z.

Named workloads (sizes as in the benchmark definitions):
  single16   16x16x16, d = 10 nm, lambda = 500 nm, eps 2.25 everywhere
  topo80     80x80x16, d = 10 nm, lambda = 542 nm, cone-filtered column design
  mixed128   128x128x32, d = 8 nm, lambda = 600 nm, Bernoulli(0.5) binary design
  batch64    64 structures, Nx,Ny in 16..64, Nz in 8..16, 2 pols x 4 angles
  serve      Poisson request stream, 8k-2M dipoles, 1-4 right-hand sides
"""
from __future__ import annotations

import math

import numpy as np

import topodda as T


def cone_filter_2d(t2d: np.ndarray, r: float) -> np.ndarray:
    """Cone filter of a 2D design: weight r - dist for dist^2 <= r^2 - 0.01,
    normalised by the weight sum, non-periodic (SpacePara.cpp:37-48,
    1241-1270; CoreStructure.cpp:64-77)."""
    nx, ny = t2d.shape
    num = np.zeros_like(t2d)
    den = np.zeros_like(t2d)
    R = int(math.ceil(r))
    for dx in range(-R, R + 1):
        for dy in range(-R, R + 1):
            if dx * dx + dy * dy > r * r - 0.01:
                continue
            w = r - math.sqrt(dx * dx + dy * dy)
            xs = slice(max(0, -dx), min(nx, nx - dx))
            ys = slice(max(0, -dy), min(ny, ny - dy))
            xd = slice(max(0, dx), min(nx, nx + dx))
            yd = slice(max(0, dy), min(ny, ny + dy))
            num[xs, ys] += w * t2d[xd, yd]
            den[xs, ys] += w
    return num / den


def extrude(t2d: np.ndarray, Nz: int) -> np.ndarray:
    """Per-site design values (3N, components equal) for a z-extruded 2D design
    in build_a_bulk order."""
    return np.repeat(t2d[:, :, None], Nz, axis=2).reshape(-1).repeat(3)


class Problem:
    """One structure with its operator inputs and right-hand-side set-up."""

    def __init__(self, name, Nx, Ny, Nz, d, lam, material, t, nback=1.0):
        self.name = name
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.d, self.lam, self.nback = d, lam, nback
        self.R = T.build_a_bulk(Nx, Ny, Nz)
        self.N = Nx * Ny * Nz
        self.K = T.wavenumber(lam, nback)
        self.material = T.material_rule(np.asarray(material, np.complex128), nback)
        self.t = np.asarray(t, np.float64)

    def rhs(self, nK, nE, E0=1.0):
        """(al, E, eps) for one incidence: DDAModel ctor (DDAModel.cpp:12-69)."""
        diel, al = T.polarizability(self.t, self.material, self.lam, self.K, self.d, nE, nK)
        E = T.incident_field(self.R, self.K, self.d, nK, nE, E0)
        return al, E, diel


def single16() -> Problem:
    N = 16 ** 3
    return Problem("single16", 16, 16, 16, 10.0, 500.0, [1.0, 2.25], np.ones(3 * N))


def topo80(seed: int = 2) -> tuple[Problem, np.ndarray]:
    rng = np.random.default_rng(seed)
    t2 = cone_filter_2d(rng.random((80, 80)), 2.0)
    p = Problem("topo80", 80, 80, 16, 10.0, 542.0, [1.0, 2.25 + 0.01j], extrude(t2, 16))
    targets = np.sort(rng.choice(p.N, 100, replace=False))
    return p, targets


def mixed128(seed: int = 3) -> Problem:
    rng = np.random.default_rng(seed)
    occ = (rng.random(128 * 128 * 32) < 0.5).astype(np.float64)
    eps = (2.0 + 0.01j) ** 2
    return Problem("mixed128", 128, 128, 32, 8.0, 600.0, [1.0, eps], occ.repeat(3))


def batch64(seed: int = 4, count: int = 64):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        Nx, Ny = int(rng.integers(16, 65)), int(rng.integers(16, 65))
        Nz = int(rng.integers(8, 17))
        t = rng.random(Nx * Ny * Nz).repeat(3)
        out.append(Problem(f"s{k}", Nx, Ny, Nz, 10.0, 500.0, [1.0, 2.25 + 0.01j], t))
    return out


def incidences():
    """2 polarisations x 4 angles in the xz plane: (nK, nE) pairs; "x" is the
    p-polarisation of nEPerpinXZ (tools.cpp:686-709), "y" is (0, 1, 0)."""
    res = []
    for deg in (0.0, 10.0, 20.0, 30.0):
        th = math.radians(deg)
        nK = np.array([math.sin(th), 0.0, math.cos(th)])
        res.append((nK, T.ne_perp_in_xz(th)))
        res.append((nK, np.array([0.0, 1.0, 0.0])))
    return res


def serve_stream(seconds: float, rate: float = 200.0, seed: int = 5):
    """Request stream: Poisson arrivals at `rate`/s for `seconds`; dipole
    count log-uniform in [8k, 2M], Nz in {8, 16, 32}, Nx = Ny =
    round(sqrt(n / Nz)); design values U[0, 1) (eps between 1.01 and
    2.25 + 0.01i); lambda U[450, 700] nm, d = 10 nm; 1-4 right-hand sides;
    half forward (z-incident plane wave, x or y polarised), half adjoint
    (complex sources at 100 seeded target dipoles, all three components).
    Returns (requests, pool): each request's design values are
    pool[off : off + 3N]."""
    rng = np.random.default_rng(seed)
    pool = rng.random(3 * 2_100_000)
    reqs = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        n = math.exp(rng.uniform(math.log(8192), math.log(2_097_152)))
        Nz = int(rng.choice([8, 16, 32]))
        Nx = max(1, int(round(math.sqrt(n / Nz))))
        N = Nx * Nx * Nz
        nrhs = int(rng.integers(1, 5))
        adjoint = bool(rng.random() < 0.5)
        incs = [(np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]) if rng.random() < 0.5 else
                 np.array([0.0, 1.0, 0.0])) for _ in range(nrhs)]
        src = None
        if adjoint:
            src = []
            for _ in range(nrhs):
                sites = rng.choice(N, 100, replace=False)
                idx = np.sort((3 * sites[:, None] + np.arange(3)[None, :]).reshape(-1)).astype(np.int32)
                src.append((idx, rng.standard_normal(300) + 1j * rng.standard_normal(300)))
        off = int(rng.integers(0, len(pool) - 3 * N + 1))
        reqs.append(dict(at=t, Nx=Nx, Nz=Nz, N=N, lam=float(rng.uniform(450.0, 700.0)), incs=incs, src=src,
                         off=off))
    return reqs, pool
