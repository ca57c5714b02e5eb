"""CPU oracle for the DDA operator and BiCGSTAB solve (Topo-DDA-forWin64).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports it.

Two layers:

* ``liboracle.so`` (oracle/dda_oracle.c): FP64 C restatement of the reference
  path -- interaction tensor (AProductCore.cpp:588-623), Get_Alpha
  (tools.cpp:617-630), incident field (DDAModel.cpp:47-51), permittivity
  interpolation and 1/alpha (DDAModel.cpp:54-63), FFT operator
  (AProductCore.cpp:12-249, 697-758; kernel.cu:74-95) and BiCGSTAB
  (DDAModel.cpp:131-260).  Wrapped by the ``Orc*`` helpers below.
* numpy references used to pin the C restatement: an O(N^2) direct sum in
  long double, dense assembly of the 3N x 3N matrix, and an ``numpy.fft``
  convolution.  They restate the same formulas independently (different
  algebraic form for the tensor: -G with the textbook dyadic Green function).

* numpy restatement of the periodic operator (AProductCore.cpp:251-536,
  573-586): an O(N^2 x images) direct sum over the image lattice with the
  textbook tensor above.
* numpy restatement of the FCD tensor (AProductCore.cpp:625-695) with the
  reference's Si / Ci tables (SiCi/SiCi.py: scipy.special.sici at
  x = 0.1, 0.2, ...) and their linear interpolation (SiCi.cpp:27-46).
* numpy restatement of the objectives PointE and IntegratedE
  (ObjDDAModel.cpp:20-177), of the reference's finite-difference adjoint
  source (EvoDDAModel.cpp:353-391) and of the closed-form source.
* numpy restatement of the design update of an extruded structure: the beta
  schedules (tools.cpp:776-798), the cone filter (SpacePara.cpp:37-48,
  1241-1270; CoreStructure.cpp:31-104) and SmoothDensity
  (FilterOption.cpp:49-60).

``oracle/_ref/libref_tools.so`` is the reference's own tools.cpp, Structure,
Space, SpacePara, FilterOption and CoreStructure compiled from its sources
(oracle/Makefile target ``ref``); ``ref_lib()`` loads it when present.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference tools library when mounted)."""
    args = ["make", "-s", "-C", HERE, "all"]
    if force:
        args.insert(1, "-B")
    subprocess.check_call(args)
    if os.path.isdir("/root/reference/Topo-DDA-forWin64"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_ldr_inter.argtypes = [ctypes.c_double] * 4 + [_dp]
        L.orc_get_alpha.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_wavenumber.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_wavenumber.restype = ctypes.c_double
        L.orc_material_rule.argtypes = [ctypes.c_int, _dp, ctypes.c_double]
        L.orc_material_rule_ex.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int]
        L.orc_einc.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_diel_al.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_local_field.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_aproduct_direct.argtypes = [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_fft1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_op_create.argtypes = [ctypes.c_int, _ip] + [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_double]
        L.orc_op_create.restype = ctypes.c_void_p
        L.orc_op_destroy.argtypes = [ctypes.c_void_p]
        L.orc_op_apply.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_bicgstab.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   _ip, _dp]
        L.orc_bicgstab.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own tools.cpp (compiled into oracle/_ref), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_tools.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_get_alpha.argtypes = [ctypes.c_double] * 5 + [_dp, _dp, _dp]
        L.ref_nE_perp_in_xz.argtypes = [ctypes.c_double, ctypes.c_double, _dp]
        L.ref_build_a_bulk.argtypes = [ctypes.c_int] * 3 + [_ip]
        L.ref_beta_update.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4
        L.ref_beta_update.restype = ctypes.c_double
        L.ref_design_update.argtypes = [ctypes.c_int] * 3 + [_dp, _dp] + [ctypes.c_double] * 3 + [_dp] * 3
        L.ref_design_update.restype = ctypes.c_int
        L.ref_structure_output.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_double, ctypes.c_char_p] + \
            [ctypes.c_int] * 2
        L.ref_structure_output.restype = ctypes.c_int
        L.ref_model_results_full.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                             ctypes.c_int, _dp, _dp]
        L.ref_model_results_full.restype = ctypes.c_int
        L.ref_spacepara_case.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ref_spacepara_case.restype = ctypes.c_int
        _REF = L
    return _REF


# ---------------------------------------------------------------------------
# Scalars and per-site quantities (C restatement)
# ---------------------------------------------------------------------------
def wavenumber(lam: float, nback: float = 1.0) -> float:
    """AProductCore.cpp:17."""
    return lib().orc_wavenumber(lam, nback)


def material_rule(mat, nback: float = 1.0, periodic: bool = False) -> np.ndarray:
    """AProductCore.cpp:19-24 (divide by nback^2, +0.01 when |eps| <= 1.01);
    periodic: the periodic constructor's strict test |eps| < 1.01
    (AProductCore.cpp:261-267)."""
    m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).copy())
    lib().orc_material_rule_ex(len(m), _ptr(m.view(np.float64)), nback, 1 if periodic else 0)
    return m


def ldr_inter(K: float, x: float, y: float, z: float) -> np.ndarray:
    """AProductCore.cpp:588-623; returns [xx, xy, xz, yy, yz, zz]."""
    out = np.zeros(6, np.complex128)
    lib().orc_ldr_inter(K, x, y, z, _ptr(out.view(np.float64)))
    return out


def get_alpha(K: float, d: float, diel: complex, nE, nK) -> complex:
    """tools.cpp:617-630."""
    out = np.zeros(2)
    e = np.ascontiguousarray(nE, dtype=np.float64)
    k = np.ascontiguousarray(nK, dtype=np.float64)
    lib().orc_get_alpha(K, d, diel.real, diel.imag, _ptr(e), _ptr(k), _ptr(out))
    return complex(out[0], out[1])


def einc(R: np.ndarray, K: float, d: float, E0: float, nK, nE) -> np.ndarray:
    """DDAModel.cpp:47-51."""
    R = np.ascontiguousarray(R, dtype=np.int32)
    N = len(R) // 3
    E = np.zeros(3 * N, np.complex128)
    lib().orc_einc(N, _ptr(R, _ip), K, d, E0, _ptr(np.ascontiguousarray(nK, np.float64)),
                   _ptr(np.ascontiguousarray(nE, np.float64)), _ptr(E.view(np.float64)))
    return E


def diel_al(diel_old: np.ndarray, mat: np.ndarray, lam: float, K: float, d: float, nE, nK):
    """DDAModel.cpp:54-63: (eps per component, 1/alpha per component)."""
    t = np.ascontiguousarray(diel_old, dtype=np.float64)
    m = np.ascontiguousarray(mat, dtype=np.complex128)
    diel = np.zeros(len(t), np.complex128)
    al = np.zeros(len(t), np.complex128)
    lib().orc_diel_al(len(t), _ptr(t), len(m), _ptr(m.view(np.float64)), lam, K, d,
                      _ptr(np.ascontiguousarray(nE, np.float64)), _ptr(np.ascontiguousarray(nK, np.float64)),
                      _ptr(diel.view(np.float64)), _ptr(al.view(np.float64)))
    return diel, al


def local_field(al: np.ndarray, P: np.ndarray, diel: np.ndarray) -> np.ndarray:
    """DDAModel.cpp:567-575."""
    out = np.zeros_like(P)
    lib().orc_local_field(len(P), _ptr(np.ascontiguousarray(al).view(np.float64)),
                          _ptr(np.ascontiguousarray(P).view(np.float64)),
                          _ptr(np.ascontiguousarray(diel).view(np.float64)), _ptr(out.view(np.float64)))
    return out


def aproduct_direct(R: np.ndarray, d: float, K: float, b: np.ndarray) -> np.ndarray:
    """O(N^2) direct sum of AProductCore::Aproduct's off-diagonal product (C)."""
    R = np.ascontiguousarray(R, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    y = np.zeros_like(b)
    lib().orc_aproduct_direct(len(R) // 3, _ptr(R, _ip), d, K, _ptr(b.view(np.float64)), _ptr(y.view(np.float64)))
    return y


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(len(y), _ptr(y.view(np.float64)), sign)
    return y


class OrcOperator:
    """FFT-based operator of the C oracle (AProductCore.cpp:12-249, 697-758)."""

    def __init__(self, R, Nx, Ny, Nz, d, K, grid=None):
        self.R = np.ascontiguousarray(R, dtype=np.int32)
        self.N = len(self.R) // 3
        Mx, My, Mz = grid if grid is not None else (2 * Nx - 1, 2 * Ny - 1, 2 * Nz - 1)
        self.grid = (Mx, My, Mz)
        self.h = lib().orc_op_create(self.N, _ptr(self.R, _ip), Nx, Ny, Nz, Mx, My, Mz, d, K)
        if not self.h:
            raise ValueError("grid smaller than 2N-1")

    def apply(self, b, al=None):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        y = np.zeros_like(b)
        alp = _ptr(np.ascontiguousarray(al, np.complex128).view(np.float64)) if al is not None else None
        lib().orc_op_apply(self.h, _ptr(b.view(np.float64)), alp, _ptr(y.view(np.float64)))
        return y

    def bicgstab(self, al, E, P0=None, max_it=10000, tol=1e-8, max_aproducts=0):
        """DDAModel.cpp:131-260.  Returns (P, status, iterations, error)."""
        al = np.ascontiguousarray(al, np.complex128)
        E = np.ascontiguousarray(E, np.complex128)
        P = np.zeros_like(E) if P0 is None else np.ascontiguousarray(P0, np.complex128).copy()
        it = ctypes.c_int(0)
        err = ctypes.c_double(0)
        st = lib().orc_bicgstab(self.h, _ptr(al.view(np.float64)), _ptr(E.view(np.float64)),
                                _ptr(P.view(np.float64)), max_it, tol, max_aproducts,
                                ctypes.byref(it), ctypes.byref(err))
        return P, st, it.value, err.value

    def close(self):
        if self.h:
            lib().orc_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Independent numpy restatements used to pin the C layer
# ---------------------------------------------------------------------------
def green_textbook(K, x, y, z, dtype=np.complex128):
    """-G(r) with G the free-space dyadic Green function written as
    e^{ikr}/r [k^2 (I - r^r^) + (ikr - 1)/r^2 (I - 3 r^r^)].  The reference's
    LDR_inter (AProductCore.cpp:588-623) is algebraically -G; this form shares
    no intermediate with it.  Returns [..., 6] as xx xy xz yy yz zz."""
    ftype = np.longdouble if dtype == np.clongdouble else np.float64
    x = np.asarray(x, ftype)
    y = np.asarray(y, ftype)
    z = np.asarray(z, ftype)
    K = ftype(K)
    r = np.sqrt(x * x + y * y + z * z)
    rs = np.where(r == 0, ftype(1), r)
    u = np.stack([x / rs, y / rs, z / rs], axis=-1)
    ph = np.exp(1j * (K * rs).astype(dtype)) if dtype != np.clongdouble else (
        np.cos(K * rs) + 1j * np.sin(K * rs)).astype(np.clongdouble)
    pre = ph / rs
    a = K * K
    bco = (1j * K * rs - 1) / (rs * rs)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    out = []
    for (i, j) in pairs:
        dij = 1.0 if i == j else 0.0
        uu = u[..., i] * u[..., j]
        g = pre * (a * (dij - uu) + bco * (dij - 3 * uu))
        out.append(np.where(r == 0, 0, -g))
    return np.stack(out, axis=-1).astype(dtype)


def assemble_dense(R, d, K, al=None):
    """Explicit 3N x 3N matrix of Aproductwithalb (DDAModel.cpp:577-586)."""
    R = np.asarray(R, np.int64).reshape(-1, 3)
    N = len(R)
    D = (R[:, None, :] - R[None, :, :]) * d
    G = green_textbook(K, D[..., 0], D[..., 1], D[..., 2])
    T = np.zeros((N, N, 3, 3), np.complex128)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    for i in range(3):
        for j in range(3):
            T[:, :, i, j] = G[..., idx[i][j]]
    A = T.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)
    if al is not None:
        A = A + np.diag(np.asarray(al, np.complex128))
    return A


def aproduct_direct_ld(R, d, K, b):
    """Long-double O(N^2) direct sum; tiny lattices only."""
    R = np.asarray(R, np.int64).reshape(-1, 3)
    D = (R[:, None, :] - R[None, :, :]).astype(np.longdouble) * np.longdouble(d)
    G = green_textbook(K, D[..., 0], D[..., 1], D[..., 2], dtype=np.clongdouble)
    b3 = np.asarray(b, np.clongdouble).reshape(-1, 3)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    y = np.zeros_like(b3)
    for i in range(3):
        for j in range(3):
            y[:, i] += np.einsum("ab,b->a", G[..., idx[i][j]], b3[:, j])
    return y.reshape(-1)


def aproduct_numpy_fft(R, Nx, Ny, Nz, d, K, b, grid=None):
    """numpy.fft convolution on grid (Mx, My, Mz) >= 2N-1 (pocketfft, FP64)."""
    Mx, My, Mz = grid if grid is not None else (2 * Nx, 2 * Ny, 2 * Nz)
    R = np.asarray(R, np.int64).reshape(-1, 3)

    def offs(M, Nn):
        i = np.arange(M)
        o = np.where(i <= Nn - 1, i, i - M).astype(np.float64)
        ok = (i <= Nn - 1) | (i >= M - (Nn - 1))
        return o, ok

    ox, kx = offs(Mx, Nx)
    oy, ky = offs(My, Ny)
    oz, kz = offs(Mz, Nz)
    X, Y, Z = np.meshgrid(ox * d, oy * d, oz * d, indexing="ij")
    mask = kx[:, None, None] & ky[None, :, None] & kz[None, None, :]
    G = green_textbook(K, X, Y, Z) * mask[..., None]
    Gh = np.fft.fftn(G, axes=(0, 1, 2))
    B = np.zeros((3, Mx, My, Mz), np.complex128)
    b3 = np.asarray(b, np.complex128).reshape(-1, 3)
    for c in range(3):
        B[c][R[:, 0], R[:, 1], R[:, 2]] = b3[:, c]
    Bh = np.fft.fftn(B, axes=(1, 2, 3))
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    C = np.zeros_like(Bh)
    for i in range(3):
        for j in range(3):
            C[i] += Gh[..., idx[i][j]] * Bh[j]
    y = np.fft.ifftn(C, axes=(1, 2, 3))
    out = np.stack([y[c][R[:, 0], R[:, 1], R[:, 2]] for c in range(3)], axis=1)
    return out.reshape(-1)


# ---------------------------------------------------------------------------
# Lattices and synthetic problems
# ---------------------------------------------------------------------------
def build_a_bulk(Nx, Ny, Nz) -> np.ndarray:
    """tools.cpp:497-510: x-major, then y, then z."""
    x, y, z = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32).reshape(-1)


def ne_perp_in_xz(theta: float) -> np.ndarray:
    """p-polarisation for k = (sin t, 0, cos t): tools.cpp:686-709 at phi = 0."""
    return np.array([math.cos(theta), 0.0, -math.sin(theta)])


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# ---------------------------------------------------------------------------
# Design update of an extruded structure (numpy restatement)
# ---------------------------------------------------------------------------
def beta_update(kind: int, x: float, x_max: float, y_min: float, y_max: float) -> float:
    """exp_update / piecewise_update / linear_update (kind 0 / 1 / 2),
    tools.cpp:776-798, as FilterOption::update_beta picks them
    (FilterOption.cpp:24-46)."""
    if kind == 0:
        return y_min + (y_max - y_min) * math.exp(x / x_max - 1.0)
    if kind == 1:
        if x <= 0.5 * x_max:
            return y_min
        if 0.5 * x_max < x <= 0.7 * x_max:
            return y_min + (y_max - y_min) / 20
        if 0.7 * x_max < x <= 0.8 * x_max:
            return y_min + (y_max - y_min) / 10
        return y_max
    return y_min + (y_max - y_min) * x / x_max


def smooth_density(x: float, beta: float, eta: float) -> float:
    """FilterOption::SmoothDensity (FilterOption.cpp:49-60); raises outside [0, 1]."""
    if 0.0 <= x <= eta:
        return eta * (math.exp(-beta * (1 - x / eta)) - (1 - x / eta) * math.exp(-beta))
    if eta < x <= 1.0:
        return (1 - eta) * (1 - math.exp(-beta * (x - eta) / (1 - eta)) + (x - eta) / (1 - eta) * math.exp(-beta)) + eta
    raise ValueError("SmoothDensity input out of range")


def cone_weights(Nx: int, Ny: int, r: float):
    """FreeWeight of an extruded Nx x Ny box (SpacePara.cpp:1241-1270): for
    every column (index x * Ny + y) the columns j with
    dx^2 + dy^2 <= r^2 - 0.01 (circlerange, SpacePara.cpp:41-48), in column
    order, with weight r - dist (calweight, SpacePara.cpp:37-39)."""
    out = []
    for xo in range(Nx):
        for yo in range(Ny):
            row = []
            for x in range(Nx):
                for y in range(Ny):
                    d2 = float(x - xo) * float(x - xo) + float(y - yo) * float(y - yo)
                    if d2 <= r * r - 0.01:
                        row.append((x * Ny + y, r - math.sqrt(d2)))
            out.append(row)
    return out


def design_update(para_origin, step, Nx: int, Ny: int, r: float, beta: float, eta: float):
    """CoreStructure::UpdateStr with a filter (CoreStructure.cpp:31-80):
    origin += step clamped to [0, 1]; filtered = sum w origin / sum w (in
    FreeWeight order); para = SmoothDensity(filtered).  Returns
    (origin, filtered, para)."""
    o = np.asarray(para_origin, np.float64).copy()
    if step is not None:
        o = o + np.asarray(step, np.float64)
        o = np.where(o >= 1, 1.0, o)
        o = np.where(o <= 0, 0.0, o)
    f = np.zeros_like(o)
    for i, row in enumerate(cone_weights(Nx, Ny, r)):
        num = den = 0.0
        for j, w in row:
            num += w * o[j]
            den += w
        f[i] = num / den
    p = np.array([smooth_density(v, beta, eta) for v in f])
    return o, f, p


def design_diel_old(para, Nx: int, Ny: int, Nz: int) -> np.ndarray:
    """diel_old of a full box in build_a_bulk order: each site takes the
    value of its column, three components each (CoreStructure.cpp:98-104)."""
    return np.repeat(np.repeat(np.asarray(para, np.float64).reshape(Nx * Ny), Nz), 3)


# ---------------------------------------------------------------------------
# Periodic operator (numpy restatement)
# ---------------------------------------------------------------------------
PERIODIC_GAMMA = 0.01  # A_dic_generator, AProductCore.cpp:574


def green_periodic(K, x, y, z, pm: int, pn: int, Lm: float, Ln: float, fcd_d=None):
    """Tensor of a structure periodic in x and y at offsets (x, y, z)
    (arrays): sum over the images m = -pm..pm, n = -pn..pn of the tensor at
    (x + m Lm, y + n Ln, z), each damped by exp(-(0.01 K r)^4) with r the
    image's distance (A_dic_generator, AProductCore.cpp:573-586; the image
    sums of AProductCore.cpp:301-312).  The tensor is green_textbook, or the
    FCD tensor of lattice spacing fcd_d when that is given (zero at r = 0).
    Returns [..., 6]."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    acc = np.zeros(np.broadcast(x, y, z).shape + (6,), np.complex128)
    for m in range(-pm, pm + 1):
        for n in range(-pn, pn + 1):
            xi, yi = x + m * Lm, y + n * Ln
            r = np.sqrt(xi * xi + yi * yi + z * z)
            damp = np.exp(-(PERIODIC_GAMMA * K * r) ** 4)
            g = green_textbook(K, xi, yi, z) if fcd_d is None else fcd_inter(K, fcd_d, xi, yi, z)
            acc += g * damp[..., None]
    return acc


def aproduct_periodic_direct(R, d, K, b, pm: int, pn: int, Lm: float, Ln: float, fcd=False):
    """y_i = sum_j G_per(r_i - r_j) b_j over all site pairs, the self pair
    included (its (0, 0) image is zero, the other images are not); fcd
    selects the FCD tensor."""
    R = np.asarray(R, np.float64).reshape(-1, 3) * d
    b3 = np.asarray(b, np.complex128).reshape(-1, 3)
    D = R[:, None, :] - R[None, :, :]
    G = green_periodic(K, D[..., 0], D[..., 1], D[..., 2], pm, pn, Lm, Ln, d if fcd else None)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    y = np.zeros_like(b3)
    for a in range(3):
        for c in range(3):
            y[:, a] += G[..., idx[a][c]] @ b3[:, c]
    return y.reshape(-1)


# ---------------------------------------------------------------------------
# Objectives and adjoint sources (numpy restatement)
# ---------------------------------------------------------------------------
def _tensor3(a6):
    a = np.asarray(a6)
    return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])


def point_e_field(R, d, K, P, r0, nK, nE, E0=1.0, fcd=False):
    """ObjPointEDDAModel (ObjDDAModel.cpp:20-97): the field at r0,
    E_ext - sum_i A(r0 - d R_i) P_i, E_ext = E0 nE exp(i K nK . r0), with the
    tensor of the C restatement (or the FCD tensor, A_dic_generator,
    AProductCore.cpp:560-571).  Returns (E, [A_i])."""
    R = np.asarray(R).reshape(-1, 3)
    P3 = np.asarray(P, np.complex128).reshape(-1, 3)
    r0 = np.asarray(r0, np.float64)
    ph = K * float(np.dot(nK, r0))
    E = E0 * np.asarray(nE, np.float64) * complex(math.cos(ph), math.sin(ph))
    As = []
    for i in range(len(R)):
        off = r0 - d * R[i]
        A = _tensor3(fcd_inter(K, d, *off) if fcd else ldr_inter(K, *off))
        As.append(A)
        E = E - A @ P3[i]
    return E, As


def field_at_points(R, d, K, P, pts, nK, nE, E0=1.0, fcd=False):
    """DDAModel::solve_E (DDAModel.cpp:530-565): E_ext - sum_i A(d pt -
    d R_i) P_i at lattice points pts (units of d), with the reference's
    incident field: component c = (x, y, z)[c]'s cosine takes
    K (nK_x q_c + nK_y y + nK_z z), its sine K nK . r (DDAModel.cpp:542-544).
    Returns (npts, 3) complex."""
    R = np.asarray(R, np.float64).reshape(-1, 3) * d
    P3 = np.asarray(P, np.complex128).reshape(-1, 3)
    pts = np.asarray(pts, np.float64).reshape(-1, 3) * d
    out = np.zeros((len(pts), 3), np.complex128)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    for q, (x, y, z) in enumerate(pts):
        D = np.array([x, y, z]) - R
        G = fcd_inter(K, d, D[:, 0], D[:, 1], D[:, 2]) if fcd else green_textbook(K, D[:, 0], D[:, 1], D[:, 2])
        s = np.array([sum(G[:, idx[a][c]] @ P3[:, c] for c in range(3)) for a in range(3)])
        sn = math.sin(K * (nK[0] * x + nK[1] * y + nK[2] * z))
        args = [nK[0] * x + nK[1] * y + nK[2] * z, nK[0] * y + nK[1] * y + nK[2] * z,
                nK[0] * z + nK[1] * y + nK[2] * z]
        Eext = np.array([E0 * nE[c] * complex(math.cos(K * args[c]), sn) for c in range(3)])
        out[q] = Eext - s
    return out


def obj_point_e(R, d, K, P, r0, nK, nE, E0=1.0) -> float:
    """ObjPointEDDAModel::GetVal: |E(r0)| (Eigen norm)."""
    return float(np.sqrt(np.sum(np.abs(point_e_field(R, d, K, P, r0, nK, nE, E0)[0]) ** 2)))


def _le_chain(lo, v, hi):
    # (lo <= v <= hi) as C++ evaluates it: (lo <= v) is 0 or 1
    return (1 if lo <= v else 0) <= hi


def integrated_e_terms(R, al, P, diel_old, pw, box, eta, beta):
    """Per entry f_i |al P|^pow of ObjIntegratedEDDAModel::SingleResponse
    (ObjDDAModel.cpp:134-158), zero outside the region test (ObjDDAModel.cpp:138)."""
    R = np.asarray(R).reshape(-1, 3)
    t = np.zeros(3 * len(R))
    fvec = np.zeros(3 * len(R))
    for i in range(len(R)):
        x, y, z = (int(v) for v in R[i])
        if _le_chain(box[0], x, box[1]) and _le_chain(box[2], y, box[3]) and _le_chain(box[4], z, box[5]):
            f = smooth_density(float(diel_old[3 * i]), beta, eta)
            fvec[3 * i:3 * i + 3] = f
            for c in range(3):
                t[3 * i + c] = f * abs(al[3 * i + c] * P[3 * i + c]) ** pw
    return t, fvec


def obj_integrated_e(R, al, P, diel_old, pw, box, eta, beta) -> float:
    """ObjIntegratedEDDAModel::GetVal: log of the sum (ObjDDAModel.cpp:160-177)."""
    return math.log(float(np.sum(integrated_e_terms(R, al, P, diel_old, pw, box, eta, beta)[0])))


def devp_fd_point_e(R, d, K, P, r0, nK, nE, E0, eps, central=False):
    """EvoDDAModel::devp (EvoDDAModel.cpp:353-391) for PointE:
    (F(P + eps e_i) - F) / eps + (F(P + i eps e_i) - F) / (i eps), each
    entry updating only its site's term (SingleResponse); central=True
    takes centred differences instead (for pinning the closed form)."""
    E, As = point_e_field(R, d, K, P, r0, nK, nE, E0)
    F = float(np.linalg.norm(E))
    out = np.zeros(len(P), np.complex128)
    for k in range(len(P)):
        i, c = divmod(k, 3)
        col = As[i][:, c]

        def Fd(delta):
            return float(np.linalg.norm(E - col * delta))
        if central:
            out[k] = (Fd(eps) - Fd(-eps)) / (2 * eps) + (Fd(1j * eps) - Fd(-1j * eps)) / (2j * eps)
        else:
            out[k] = (Fd(eps) - F) / eps + (Fd(1j * eps) - F) / (1j * eps)
    return out


def devp_point_e(R, d, K, P, r0, nK, nE, E0=1.0):
    """Closed form of the PointE source: -(1/F) sum_a A_ac conj(E_a)."""
    E, As = point_e_field(R, d, K, P, r0, nK, nE, E0)
    F = float(np.linalg.norm(E))
    return np.concatenate([-(A.T @ np.conj(E)) / F for A in As])


def devp_fd_integrated_e(R, al, P, diel_old, pw, box, eta, beta, eps, central=False):
    """EvoDDAModel::devp for IntegratedE (F = log E_int)."""
    t, fvec = integrated_e_terms(R, al, P, diel_old, pw, box, eta, beta)
    S = float(np.sum(t))
    out = np.zeros(len(P), np.complex128)
    for k in range(len(P)):
        def Fd(delta):
            return math.log(S - t[k] + fvec[k] * abs(al[k] * (P[k] + delta)) ** pw)
        if central:
            out[k] = (Fd(eps) - Fd(-eps)) / (2 * eps) + (Fd(1j * eps) - Fd(-1j * eps)) / (2j * eps)
        else:
            F = math.log(S)
            out[k] = (Fd(eps) - F) / eps + (Fd(1j * eps) - F) / (1j * eps)
    return out


def devp_integrated_e(R, al, P, diel_old, pw, box, eta, beta):
    """Closed form of the IntegratedE source: pow f |al|^pow |P|^(pow-2) conj(P) / E_int."""
    t, fvec = integrated_e_terms(R, al, P, diel_old, pw, box, eta, beta)
    S = float(np.sum(t))
    aP = np.abs(P)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(aP > 0, pw * fvec * np.abs(al) ** pw * aP ** (pw - 2.0) * np.conj(P), 0)
    return g / S


# ---------------------------------------------------------------------------
# FCD tensor (numpy restatement)
# ---------------------------------------------------------------------------
SICI_STEP = 0.1  # SiCi/SiCi.py:10


def sici_table(x):
    """SiCi::get_Si / get_Ci (SiCi.cpp:27-46): linear interpolation between
    the nodes floor(x / 0.1) 0.1 and ceil(x / 0.1) 0.1 of tables holding
    scipy.special.sici at the nodes (SiCi/SiCi.py); on a node, the node's
    value.  Scalars or arrays."""
    from scipy.special import sici
    x = np.asarray(x, np.float64)
    a1 = np.floor(x / SICI_STEP) * SICI_STEP
    a2 = np.ceil(x / SICI_STEP) * SICI_STEP
    s1, c1 = sici(a1)
    s2, c2 = sici(a2)
    on = a1 == a2
    den = np.where(on, 1.0, a2 - a1)
    si = np.where(on, s1, s1 + ((s2 - s1) / den) * (x - a1))
    ci = np.where(on, c1, c1 + ((c2 - c1) / den) * (x - a1))
    if si.ndim == 0:
        return float(si), float(ci)
    return si, ci


def fcd_radial(K, d, r, sici=None):
    """The radial functions of AProductCore::FCD_inter (AProductCore.cpp:
    625-695): the filtered scalar Green function gf(r), its first and second
    r-derivatives gf1, gf2, and the filter kernel hr (r > 0, scalars or
    arrays).  `sici` (x -> (Si, Ci)) defaults to the reference's tables."""
    sici = sici or sici_table
    r = np.asarray(r, np.float64)
    Kf = math.pi / d
    up, dn = (Kf + K) * r, (Kf - K) * r
    Sz, Cz = sici(up)
    Sf, Cf = sici(dn)
    Cz1, Cf1 = np.cos(up) / r, np.cos(dn) / r
    Sz1, Sf1 = np.sin(up) / r, np.sin(dn) / r
    Cz2 = (-up * np.sin(up) - np.cos(up)) / r ** 2
    Cf2 = (-dn * np.sin(dn) - np.cos(dn)) / r ** 2
    Sz2 = (up * np.cos(up) - np.sin(up)) / r ** 2
    Sf2 = (dn * np.cos(dn) - np.sin(dn)) / r ** 2
    kr = K * r
    sk, ck = np.sin(kr), np.cos(kr)
    CC = math.pi * 1j + Cf - Cz
    C1C1, C2C2 = Cf1 - Cz1, Cf2 - Cz2
    SS, S1S1, S2S2 = Sz + Sf, Sz1 + Sf1, Sz2 + Sf2
    gf = (sk * CC + ck * SS) / (math.pi * r)
    c1 = (kr * ck - sk) * CC - (kr * sk + ck) * SS + r * sk * C1C1 + r * ck * S1S1
    c2 = (-K * kr * sk) * CC - (K * kr * ck) * SS + (2 * kr * ck) * C1C1 - (2 * kr * sk) * S1S1 + r * sk * C2C2 \
        + r * ck * S2S2
    gf1 = c1 / (math.pi * r ** 2)
    gf2 = (r * c2 - 2.0 * c1) / (math.pi * r ** 3)
    hr = (np.sin(Kf * r) - Kf * r * np.cos(Kf * r)) / (2.0 * math.pi ** 2 * r ** 3)
    return gf, gf1, gf2, hr


def fcd_inter(K, d, x, y, z, sici=None):
    """AProductCore::FCD_inter (AProductCore.cpp:625-695): [..., 6] as
    xx xy xz yy yz zz of -(K^2 gf I + grad grad gf + (4 pi / 3) hr I), zero
    at r = 0."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    r = np.sqrt(x * x + y * y + z * z)
    zero = r == 0.0
    rs = np.where(zero, 1.0, r)
    gf, gf1, gf2, hr = fcd_radial(K, d, rs, sici)
    t1 = K * K * gf + gf1 / rs + (4.0 * math.pi / 3.0) * hr
    t2 = (gf2 - gf1 / rs) / rs ** 2
    out = -np.stack([t1 + x * x * t2, x * y * t2, z * x * t2, t1 + y * y * t2, y * z * t2, t1 + z * z * t2],
                    axis=-1)
    return np.where(zero[..., None], 0.0, out)


def aproduct_fcd_direct(R, d, K, b):
    """O(N^2) direct sum of the off-diagonal product with the FCD tensor."""
    R = np.asarray(R, np.float64).reshape(-1, 3) * d
    b3 = np.asarray(b, np.complex128).reshape(-1, 3)
    D = R[:, None, :] - R[None, :, :]
    G = fcd_inter(K, d, D[..., 0], D[..., 1], D[..., 2])
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    y = np.zeros_like(b3)
    for a in range(3):
        for c in range(3):
            y[:, a] += G[..., idx[a][c]] @ b3[:, c]
    return y.reshape(-1)


