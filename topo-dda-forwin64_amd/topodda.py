"""ctypes binding of libtopodda_hip.so (C ABI declared in include/topodda.h).

Used by the tests, bench.py and __graft_entry__.  Loading fails loudly when the
HIP library has not been built: there is no CPU fallback.

Class names follow the reference's call surface (Topo-DDA-forWin64):
``AProductCore`` (AProductCore.h:51-80) owns the device operator and
``DDAModel`` (DDAModel.h:34-74) owns device-resident solver state.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("TDDA_LIB", os.path.join(HERE, "libtopodda_hip.so"))

C128 = 1
C64 = 2
VEC_AL, VEC_E, VEC_P, VEC_Y = 0, 1, 2, 3
MODE_FP64, MODE_MIXED = 0, 1
CONVERGED, BREAKDOWN, MAXIT = 1, 2, 3
PRECOND_NONE, PRECOND_JACOBI = 0, 1

EXPORTS = [
    "tdda_last_error", "tdda_version", "tdda_ctx_create", "tdda_ctx_destroy", "tdda_ctx_sync",
    "tdda_op_create", "tdda_op_destroy", "tdda_op_info", "tdda_op_plan", "tdda_aproduct", "tdda_green_tensor",
    "tdda_model_create", "tdda_model_destroy", "tdda_model_set", "tdda_model_get", "tdda_model_apply",
    "tdda_solve_opts_default", "tdda_model_solve", "tdda_model_get_entries", "tdda_model_set_entries",
    "tdda_model_time_passes", "tdda_model_time_pass", "tdda_material_rule", "tdda_material_rule_periodic", "tdda_get_alpha", "tdda_incident_field",
    "tdda_polarizability", "tdda_local_field", "tdda_batch_create", "tdda_batch_destroy", "tdda_batch_size",
    "tdda_batch_solve", "tdda_batch_apply", "tdda_batch_time_passes", "tdda_batch_time_pass", "tdda_server_opts_default",
    "tdda_server_create", "tdda_server_destroy", "tdda_server_submit", "tdda_server_poll", "tdda_server_wait",
    "tdda_server_result", "tdda_server_get_stats", "tdda_design_create", "tdda_design_create_ex", "tdda_design_step", "tdda_design_get",
    "tdda_design_destroy", "tdda_beta_update", "tdda_op_create_periodic", "tdda_objective_point_e",
    "tdda_objective_integrated_e", "tdda_op_autotune", "tdda_op_autotune_ex", "tdda_op_tiles", "tdda_op_xpass", "tdda_op_create_ex",
    "tdda_green_tensor_fcd", "tdda_field_at_points", "tdda_write_structure", "tdda_write_model_results", "tdda_model_setup", "tdda_aproduct_noalloc",
]


class TddaError(RuntimeError):
    pass


class OpDesc(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("R", ctypes.POINTER(ctypes.c_int)), ("d", ctypes.c_double), ("lam", ctypes.c_double),
                ("nback", ctypes.c_double), ("Mx", ctypes.c_int), ("My", ctypes.c_int), ("Mz", ctypes.c_int),
                ("precisions", ctypes.c_int)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("max_it", ctypes.c_int), ("tol", ctypes.c_double), ("mode", ctypes.c_int),
                ("x0_zero", ctypes.c_int), ("inner_tol", ctypes.c_double), ("inner_max_it", ctypes.c_int),
                ("max_outer", ctypes.c_int), ("check_every", ctypes.c_int), ("true_residual", ctypes.c_int),
                ("precond", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("error", ctypes.c_double),
                ("true_residual", ctypes.c_double), ("outer", ctypes.c_int)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("max_rhs", ctypes.c_int), ("solve", SolveOpts)]


class Request(ctypes.Structure):
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("op", OpDesc), ("diel_old", _dp), ("nmat", ctypes.c_int), ("material", _dp),
                ("nrhs", ctypes.c_int), ("nK", _dp), ("nE", _dp), ("E0", _dp), ("nsrc", ctypes.c_int),
                ("src_idx", ctypes.POINTER(ctypes.c_int)), ("src_vals", _dp)]


class RequestTimes(ctypes.Structure):
    _fields_ = [("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("footprint_bytes", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
                ("ticks", ctypes.c_longlong), ("queued", ctypes.c_int), ("active_requests", ctypes.c_int),
                ("active_rhs", ctypes.c_int), ("used_bytes", ctypes.c_double), ("peak_bytes", ctypes.c_double),
                ("budget_bytes", ctypes.c_double), ("setup_s", ctypes.c_double), ("tick_s", ctypes.c_double),
                ("release_s", ctypes.c_double)]


_lib = None


def lib():
    """Loads the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise TddaError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.tdda_last_error.restype = ctypes.c_char_p
        L.tdda_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.tdda_ctx_destroy.argtypes = [vp]
        L.tdda_ctx_sync.argtypes = [vp]
        L.tdda_op_create.argtypes = [vp, ctypes.POINTER(OpDesc), ctypes.POINTER(vp)]
        L.tdda_op_create_periodic.argtypes = [vp, ctypes.POINTER(OpDesc), ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.POINTER(vp)]
        L.tdda_op_destroy.argtypes = [vp]
        L.tdda_op_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
        L.tdda_op_plan.argtypes = [ctypes.POINTER(OpDesc), ctypes.POINTER(ctypes.c_int)]
        L.tdda_aproduct.argtypes = [vp, dp, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int]
        L.tdda_green_tensor.argtypes = [ctypes.c_double] * 4 + [dp]
        L.tdda_model_create.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.tdda_model_destroy.argtypes = [vp]
        L.tdda_model_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_model_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_model_apply.argtypes = [vp, ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.tdda_model_get_entries.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp]
        L.tdda_model_set_entries.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp]
        L.tdda_model_time_passes.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_model_time_pass.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_material_rule.argtypes = [ctypes.c_int, dp, ctypes.c_double]
        L.tdda_material_rule_periodic.argtypes = [ctypes.c_int, dp, ctypes.c_double]
        L.tdda_get_alpha.argtypes = [ctypes.c_double] * 5 + [dp, dp, dp]
        L.tdda_incident_field.argtypes = [ctypes.c_int, ip] + [ctypes.c_double] * 3 + [dp, dp, dp]
        L.tdda_polarizability.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp] + [ctypes.c_double] * 3 + [dp] * 4
        L.tdda_local_field.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.tdda_solve_opts_default.argtypes = [ctypes.POINTER(SolveOpts)]
        L.tdda_solve_opts_default.restype = None
        L.tdda_model_solve.argtypes = [vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
        L.tdda_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.tdda_batch_destroy.argtypes = [vp]
        L.tdda_batch_size.argtypes = [vp]
        L.tdda_batch_solve.argtypes = [vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
        L.tdda_batch_apply.argtypes = [vp, ctypes.c_int]
        L.tdda_batch_time_passes.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_batch_time_pass.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_server_opts_default.argtypes = [ctypes.POINTER(ServerOpts)]
        L.tdda_server_opts_default.restype = None
        L.tdda_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]
        L.tdda_server_destroy.argtypes = [vp]
        L.tdda_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]
        L.tdda_server_poll.argtypes = [vp, ctypes.c_longlong]
        L.tdda_server_wait.argtypes = [vp, ctypes.c_longlong, ctypes.c_double]
        L.tdda_server_result.argtypes = [vp, ctypes.c_longlong, dp, ctypes.POINTER(SolveInfo),
                                         ctypes.POINTER(RequestTimes)]
        L.tdda_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.tdda_model_setup.argtypes = [vp, ctypes.POINTER(Request)]
        L.tdda_design_create.argtypes = [vp, dp, ctypes.c_double, dp, ctypes.c_int, dp, ctypes.POINTER(vp)]
        L.tdda_design_create_ex.argtypes = [vp, ctypes.POINTER(DesignDesc), dp, ctypes.c_int, dp, ctypes.POINTER(vp)]
        L.tdda_design_step.argtypes = [vp, dp, ctypes.c_double, ctypes.c_double]
        L.tdda_design_get.argtypes = [vp, dp, dp, dp]
        L.tdda_design_destroy.argtypes = [vp]
        L.tdda_design_destroy.restype = None
        L.tdda_beta_update.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4
        L.tdda_beta_update.restype = ctypes.c_double
        L.tdda_op_create_ex.argtypes = [vp, ctypes.POINTER(OpDesc), ctypes.POINTER(OpExt), ctypes.POINTER(vp)]
        L.tdda_green_tensor_fcd.argtypes = [ctypes.c_double] * 5 + [dp]
        L.tdda_field_at_points.argtypes = [vp, ctypes.c_int, ip, ctypes.c_int, dp, dp, ctypes.c_double, dp]
        L.tdda_write_structure.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 4 + [ip, dp, ctypes.c_double,
                                                                                 ctypes.c_int]
        L.tdda_write_model_results.argtypes = [ctypes.c_char_p, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ip,
                                               ctypes.c_int, dp, dp]
        L.tdda_op_autotune.argtypes = [vp, ctypes.c_int]
        L.tdda_op_autotune_ex.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tdda_op_tiles.argtypes = [vp, ctypes.c_int, ip]
        L.tdda_op_xpass.argtypes = [vp, ctypes.c_int]
        L.tdda_objective_point_e.argtypes = [vp, ctypes.c_int, dp, dp, dp, ctypes.c_double, dp, vp, ctypes.c_int]
        L.tdda_objective_integrated_e.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, ctypes.c_double,
                                                  ctypes.c_double, dp, dp, vp, ctypes.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise TddaError(f"topodda error {rc}: {lib().tdda_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        _check(lib().tdda_ctx_create(device, ctypes.byref(self.h)))

    def sync(self):
        _check(lib().tdda_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().tdda_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def green_tensor(K, x, y, z):
    """AProductCore::LDR_inter (AProductCore.cpp:588-623) through the C ABI."""
    out = np.zeros(12)
    _check(lib().tdda_green_tensor(K, x, y, z, _dptr(out)))
    return out.view(np.complex128)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def wavenumber(lam, nback=1.0):
    """K = 2 pi / (lam / nback) (AProductCore.cpp:17)."""
    return 2 * math.pi / (lam / nback)


def material_rule(material, nback=1.0, periodic=False):
    """AProductCore.cpp:19-24 (+0.01 when |eps| <= 1.01); periodic: the
    periodic constructor's |eps| < 1.01 (AProductCore.cpp:261-267)."""
    m = _c(material).copy()
    fn = lib().tdda_material_rule_periodic if periodic else lib().tdda_material_rule
    _check(fn(len(m), _dptr(m.view(np.float64)), nback))
    return m


def get_alpha(lam, K, d, diel, nE, nK):
    """Get_Alpha (tools.cpp:617-630)."""
    out = np.zeros(2)
    _check(lib().tdda_get_alpha(lam, K, d, complex(diel).real, complex(diel).imag, _dptr(_f(nE)), _dptr(_f(nK)),
                                _dptr(out)))
    return complex(out[0], out[1])


def incident_field(R, K, d, nK, nE, E0=1.0):
    """DDAModel.cpp:47-51."""
    R = np.ascontiguousarray(R, dtype=np.int32)
    N = len(R) // 3
    E = np.zeros(3 * N, np.complex128)
    _check(lib().tdda_incident_field(N, R.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), K, d, E0, _dptr(_f(nK)),
                                     _dptr(_f(nE)), _dptr(E.view(np.float64))))
    return E


def polarizability(diel_old, material, lam, K, d, nE, nK):
    """(eps, 1/alpha) per component (DDAModel.cpp:54-63)."""
    t = _f(diel_old)
    m = _c(material)
    diel = np.zeros(len(t), np.complex128)
    al = np.zeros(len(t), np.complex128)
    _check(lib().tdda_polarizability(len(t), _dptr(t), len(m), _dptr(m.view(np.float64)), lam, K, d, _dptr(_f(nE)),
                                     _dptr(_f(nK)), _dptr(diel.view(np.float64)), _dptr(al.view(np.float64))))
    return diel, al


def local_field(al, P, diel):
    """DDAModel.cpp:567-575."""
    al, P, diel = _c(al), _c(P), _c(diel)
    out = np.zeros_like(P)
    _check(lib().tdda_local_field(len(P), _dptr(al.view(np.float64)), _dptr(P.view(np.float64)),
                                  _dptr(diel.view(np.float64)), _dptr(out.view(np.float64))))
    return out


def build_a_bulk(Nx, Ny, Nz):
    """Lattice sites in x-major, then y, then z order (tools.cpp:497-510)."""
    x, y, z = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32).reshape(-1)


def ne_perp_in_xz(theta):
    """p-polarisation perpendicular to k = (sin t, 0, cos t) in the xz plane
    with non-negative x component (nEPerpinXZ, tools.cpp:686-709, phi = 0)."""
    return np.array([math.cos(theta), 0.0, -math.sin(theta)])


class OpExt(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("max_m", ctypes.c_int), ("max_n", ctypes.c_int),
                ("Lm", ctypes.c_double), ("Ln", ctypes.c_double)]


METHOD_LDR, METHOD_FCD = 0, 1


def green_tensor_fcd(K, d, x, y, z):
    """AProductCore::FCD_inter (AProductCore.cpp:625-695): [xx xy xz yy yz zz]."""
    out = np.zeros(12)
    _check(lib().tdda_green_tensor_fcd(K, d, x, y, z, _dptr(out)))
    return out.view(np.complex128)


def op_plan(Nx, Ny, Nz, d=10.0, lam=500.0, nback=1.0, N=None, grid=None):
    """FFT grid an operator of this box would use, after the checks of
    tdda_op_create that need no device (tdda_op_plan); raises TddaError
    where tdda_op_create would refuse the description."""
    g = grid or (0, 0, 0)
    desc = OpDesc(Nx * Ny * Nz if N is None else N, Nx, Ny, Nz, None, d, lam, nback, g[0], g[1], g[2], C128 | C64)
    g3 = (ctypes.c_int * 3)()
    _check(lib().tdda_op_plan(ctypes.byref(desc), g3))
    return tuple(g3)


class AProductCore:
    """Device operator: AProductCore (AProductCore.cpp:12-249, 697-758).
    periodic=(max_m, max_n, Lm, Ln) builds the periodic operator
    (AProductCore.cpp:251-536): images m = -max_m..max_m, n = -max_n..max_n
    shifted by (m Lm, n Ln, 0), each damped by exp(-(0.01 K r)^4).
    method: "LDR" (AProductCore.cpp:588-623) or "FCD" (625-695)."""

    def __init__(self, ctx: Context, R, Nx, Ny, Nz, d, lam, nback=1.0, grid=None, precisions=C128 | C64,
                 periodic=None, method="LDR"):
        self.ctx = ctx
        # R = None: the full box in build_a_bulk order
        self.R = None if R is None else np.ascontiguousarray(R, dtype=np.int32)
        self.N = Nx * Ny * Nz if R is None else len(self.R) // 3
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.d, self.lam, self.nback = d, lam, nback
        g = grid or (0, 0, 0)
        Rp = None if R is None else self.R.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        desc = OpDesc(self.N, Nx, Ny, Nz, Rp, d, lam, nback, g[0], g[1], g[2], precisions)
        self.h = ctypes.c_void_p()
        if method not in ("LDR", "FCD"):
            raise TddaError(f'"{method}" is not a valid polarizability calculation method')
        if periodic is None and method == "LDR":
            _check(lib().tdda_op_create(ctx.h, ctypes.byref(desc), ctypes.byref(self.h)))
        else:
            pm, pn, Lm, Ln = periodic if periodic is not None else (-1, -1, 0.0, 0.0)
            ext = OpExt(METHOD_FCD if method == "FCD" else METHOD_LDR, int(pm), int(pn), float(Lm), float(Ln))
            _check(lib().tdda_op_create_ex(ctx.h, ctypes.byref(desc), ctypes.byref(ext), ctypes.byref(self.h)))
        g3 = (ctypes.c_int * 3)()
        K = ctypes.c_double()
        _check(lib().tdda_op_info(self.h, g3, ctypes.byref(K)))
        self.grid = tuple(g3)
        self.K = K.value

    def autotune(self, reps=20, nrhs=1):
        """Tile autotuner (tdda_op_autotune_ex) on a scratch model of nrhs
        right-hand sides: affects models created afterwards."""
        _check(lib().tdda_op_autotune_ex(self.h, int(reps), int(nrhs)))

    def tiles(self, precision=C128):
        """(z pencils P5, z pencils P1, y columns, x columns) per tile."""
        out = (ctypes.c_int * 4)()
        _check(lib().tdda_op_tiles(self.h, precision, out))
        return tuple(out)

    def xpass(self, precision=C128):
        """x-pass kernel (tdda_op_xpass bit 0): "register" or "lds"."""
        rc = lib().tdda_op_xpass(self.h, precision)
        _check(min(rc, 0))
        return "register" if rc & 1 else "lds"

    def ypass(self, precision=C128):
        """y-pass kernels (tdda_op_xpass bit 1): "register" or "lds"."""
        rc = lib().tdda_op_xpass(self.h, precision)
        _check(min(rc, 0))
        return "register" if rc & 2 else "lds"

    def zy_fused(self, precision=C128):
        """True when the z and y passes run fused (tdda_op_xpass bit 2): one
        workgroup per x slab keeps the z-transformed slab in LDS."""
        rc = lib().tdda_op_xpass(self.h, precision)
        _check(min(rc, 0))
        return bool(rc & 4)

    def Aproduct(self, b, al=None, precision=C128):
        """y = A_offdiag b (+ al o b); b is (3N,) or (nrhs, 3N) complex128."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        y = np.zeros_like(b)
        alp = None
        per = 0
        if al is not None:
            al = np.ascontiguousarray(al, dtype=np.complex128)
            per = 1 if al.ndim == 2 else 0
            alp = _dptr(al.view(np.float64))
        _check(lib().tdda_aproduct(self.h, _dptr(b.view(np.float64)), _dptr(y.view(np.float64)), nrhs, alp, per,
                                   precision))
        return y

    def close(self):
        if self.h:
            lib().tdda_op_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _infos(info):
    return [dict(status=i.status, iterations=i.iterations, error=i.error, true_residual=i.true_residual,
                 outer=i.outer) for i in info]


def solve_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    lib().tdda_solve_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class DeviceModel:
    """Device-resident state of nrhs DDAModel right-hand sides (al, E, P)."""

    def __init__(self, core: AProductCore, nrhs: int = 1):
        self.core = core
        self.nrhs = nrhs
        self.n3 = 3 * core.N
        self.h = ctypes.c_void_p()
        _check(lib().tdda_model_create(core.h, nrhs, ctypes.byref(self.h)))

    def set(self, which, arr, rhs=-1):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        need = self.n3 * (self.nrhs if rhs < 0 else 1)
        if a.size != need:
            raise ValueError(f"expected {need} complex values, got {a.size}")
        _check(lib().tdda_model_set(self.h, which, rhs, _dptr(a.view(np.float64))))

    def get(self, which, rhs=-1):
        n = self.n3 * (self.nrhs if rhs < 0 else 1)
        out = np.zeros(n, np.complex128)
        _check(lib().tdda_model_get(self.h, which, rhs, _dptr(out.view(np.float64))))
        return out.reshape(self.nrhs, self.n3) if rhs < 0 else out

    def setup(self, t, material, incidences, E0=None, R=None, sources=None):
        """al and E of every right-hand side built on the device from the
        design values t (3N), the material list (after material_rule) and
        the incidences [(nK, nE), ...] (tdda_model_setup: the DDAModel
        constructor, DDAModel.cpp:12-69); sources: sparse right-hand sides
        as in Server.submit.  P = 0 afterwards."""
        c = self.core
        q, keep = _request(c.Nx, c.Ny, c.Nz, c.d, c.lam, t, material, incidences, E0,
                           c.R if R is None else R, sources, c.nback)
        _check(lib().tdda_model_setup(self.h, ctypes.byref(q)))
        del keep

    def get_entries(self, which, rhs, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros(len(idx), np.complex128)
        _check(lib().tdda_model_get_entries(self.h, which, rhs, len(idx), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            _dptr(out.view(np.float64))))
        return out

    def set_entries(self, which, rhs, idx, vals):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        _check(lib().tdda_model_set_entries(self.h, which, rhs, len(idx), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            _dptr(v.view(np.float64))))

    def time_passes(self, precision=C128, reps=10):
        out = np.zeros(5)
        _check(lib().tdda_model_time_passes(self.h, precision, reps, _dptr(out)))
        return out

    def time_pass(self, pass_index, precision=C128, reps=50):
        """Mean ms of one pass kind launched reps times back to back."""
        out = ctypes.c_double()
        _check(lib().tdda_model_time_pass(self.h, precision, pass_index, reps, ctypes.byref(out)))
        return out.value

    def apply(self, precision=C128):
        _check(lib().tdda_model_apply(self.h, precision))
        return self.get(VEC_Y)

    def solve(self, **kw):
        o = solve_opts(**kw)
        info = (SolveInfo * self.nrhs)()
        _check(lib().tdda_model_solve(self.h, ctypes.byref(o), info))
        return _infos(info)

    def close(self):
        if self.h:
            lib().tdda_model_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Right-hand sides of several DeviceModels (any structures, grids and
    wavelengths on one context) solved together: each operator pass is one
    launch over all of them.  Results equal the models' solves alone."""

    def __init__(self, models):
        self.models = list(models)
        arr = (ctypes.c_void_p * len(self.models))(*[m.h.value for m in self.models])
        self.h = ctypes.c_void_p()
        _check(lib().tdda_batch_create(arr, len(self.models), ctypes.byref(self.h)))
        self.size = lib().tdda_batch_size(self.h)

    def solve(self, **kw):
        o = solve_opts(**kw)
        info = (SolveInfo * self.size)()
        _check(lib().tdda_batch_solve(self.h, ctypes.byref(o), info))
        return _infos(info)

    def apply(self, precision=C128):
        _check(lib().tdda_batch_apply(self.h, precision))

    def time_passes(self, precision=C128, reps=10):
        out = np.zeros(5)
        _check(lib().tdda_batch_time_passes(self.h, precision, reps, _dptr(out)))
        return out

    def time_pass(self, pass_index, precision=C128, reps=20):
        """Mean ms of all launches of one pass kind, repeated reps times back to back."""
        out = ctypes.c_double()
        _check(lib().tdda_batch_time_pass(self.h, precision, pass_index, reps, ctypes.byref(out)))
        return out.value

    def close(self):
        if self.h:
            lib().tdda_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def objective_point_e(model, rhs, r0, nK, nE, E0=1.0, source=None, source_rhs=0):
    """ObjPointEDDAModel value (ObjDDAModel.cpp:20-97) of right-hand side rhs;
    with source=(a DeviceModel on the same operator), its closed-form adjoint
    source 2 dF/dP (EvoDDAModel::devp, EvoDDAModel.cpp:353-391) is written to
    that model's E of source_rhs."""
    v = ctypes.c_double()
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r0, nK, nE)]
    _check(lib().tdda_objective_point_e(model.h, rhs, _dptr(a[0]), _dptr(a[1]), _dptr(a[2]), float(E0),
                                        ctypes.byref(v), source.h if source is not None else None, source_rhs))
    return v.value


def field_at_points(model, rhs, pts, nK, nE, E0=1.0):
    """DDAModel::solve_E (DDAModel.cpp:530-565) on the device: the field at
    lattice points pts ((npts, 3) ints, units of d) from right-hand side
    rhs's P, with the reference's incident field.  Returns (npts, 3)."""
    p = np.ascontiguousarray(np.asarray(pts).reshape(-1, 3), dtype=np.int32)
    out = np.zeros((len(p), 6))
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (nK, nE)]
    _check(lib().tdda_field_at_points(model.h, rhs, p.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(p),
                                      _dptr(a[0]), _dptr(a[1]), float(E0), _dptr(out)))
    return out.view(np.complex128)


RESULTS_FULL, RESULTS_E_P, RESULTS_E = 0, 1, 2


def write_structure(path, Nx, Ny, Nz, R, diel_old, d, full=True):
    """CoreStructure::output_to_file (CoreStructure.cpp:155-190): the
    reference's bytes; full=False writes diel_old only."""
    r = np.ascontiguousarray(R, dtype=np.int32)
    dl = np.ascontiguousarray(diel_old, dtype=np.float64)
    _check(lib().tdda_write_structure(os.fsencode(path), Nx, Ny, Nz, len(r) // 3,
                                      r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(dl), float(d), int(full)))


def write_model_results(path, layout, P=None, EResult=None, RResult=None, nK=(0, 0, 1), nE0=(1, 0, 0)):
    """DDAModel::output_to_file (DDAModel.cpp:589-697) with layout
    RESULTS_FULL, RESULTS_E_P or RESULTS_E."""
    Pv = np.ascontiguousarray(P if P is not None else np.zeros(0), dtype=np.complex128)
    Ev = np.ascontiguousarray(EResult if EResult is not None else np.zeros(0), dtype=np.complex128)
    Rv = np.ascontiguousarray(RResult if RResult is not None else np.zeros(0), dtype=np.int32)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (nK, nE0)]
    _check(lib().tdda_write_model_results(os.fsencode(path), int(layout), _dptr(Pv.view(np.float64)), len(Pv),
                                          _dptr(Ev.view(np.float64)), len(Ev),
                                          Rv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(Rv), _dptr(a[0]),
                                          _dptr(a[1])))


def objective_integrated_e(model, rhs, pow_num, box, eta, beta, diel_old, source=None, source_rhs=0):
    """ObjIntegratedEDDAModel value (ObjDDAModel.cpp:100-177); box = (xMin,
    xMax, yMin, yMax, zMin, zMax); source as for objective_point_e."""
    v = ctypes.c_double()
    b6 = np.ascontiguousarray(box, dtype=np.int32)
    dl = np.ascontiguousarray(diel_old, dtype=np.float64)
    _check(lib().tdda_objective_integrated_e(model.h, rhs, int(pow_num), b6.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             float(eta), float(beta), _dptr(dl), ctypes.byref(v),
                                             source.h if source is not None else None, source_rhs))
    return v.value


BETA_EXP, BETA_PIECEWISE, BETA_LINEAR = 0, 1, 2


def beta_update(kind, it, max_it, beta_min, beta_max):
    """FilterOption::update_beta (FilterOption.cpp:24-46)."""
    return lib().tdda_beta_update(kind, float(it), float(max_it), float(beta_min), float(beta_max))


class DesignDesc(ctypes.Structure):
    _ip = ctypes.POINTER(ctypes.c_int)
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("npara", ctypes.c_int), ("nfree", ctypes.c_int), ("free_index", _ip), ("site_param", _ip),
                ("para", _dp), ("filter", ctypes.c_int), ("weight_start", _ip), ("weight_pos", _ip),
                ("weight_val", _dp)]


class Design:
    """Device-resident design kept in HBM.  step() is
    CoreStructure::UpdateStr (CoreStructure.cpp:31-104: add, clamp, cone
    filter, SmoothDensity) followed by DDAModel::UpdateAlpha
    (DDAModel.cpp:470-484) for every right-hand side of the model.
    incidences: one (nK, nE) pair per right-hand side of the model.

    Design(model, para_origin, rfilter, material, incidences): an extruded
    full box, one parameter per (x, y) column (index x * Ny + y).
    Design.general(...): any SpacePara parameterization."""

    def __init__(self, model: "DeviceModel", para_origin, rfilter, material, incidences, _desc=None):
        self.model = model
        mat = np.ascontiguousarray(material, dtype=np.complex128)
        inc = np.ascontiguousarray([np.concatenate([np.asarray(k, float), np.asarray(e, float)])
                                    for k, e in incidences], dtype=np.float64).reshape(-1)
        if len(inc) != 6 * model.nrhs:
            raise TddaError("one (nK, nE) incidence per right-hand side of the model")
        self.h = ctypes.c_void_p()
        if _desc is None:
            po = np.ascontiguousarray(para_origin, dtype=np.float64)
            self.npara = self.nfree = len(po)
            _check(lib().tdda_design_create(model.h, _dptr(po), float(rfilter), _dptr(mat.view(np.float64)),
                                            len(mat), _dptr(inc), ctypes.byref(self.h)))
        else:
            d, keep = _desc
            self.npara, self.nfree = d.npara, d.nfree
            _check(lib().tdda_design_create_ex(model.h, ctypes.byref(d), _dptr(mat.view(np.float64)), len(mat),
                                               _dptr(inc), ctypes.byref(self.h)))
            del keep

    @classmethod
    def general(cls, model, site_param, para, free_index, material, incidences, weights=None):
        """site_param: the parameter of each site of the model's operator
        (SpacePara geometryPara); para: all parameter values; free_index: the
        free parameters' indices in para; weights: None (no filter) or, per
        free parameter, [(para index, weight), ...] in SpacePara's FreeWeight
        order (tdda_design_create_ex)."""
        ip = ctypes.POINTER(ctypes.c_int)
        sp = np.ascontiguousarray(site_param, dtype=np.int32)
        pa = np.ascontiguousarray(para, dtype=np.float64)
        fi = np.ascontiguousarray(free_index, dtype=np.int32)
        d = DesignDesc()
        d.npara, d.nfree = len(pa), len(fi)
        d.free_index = fi.ctypes.data_as(ip)
        d.site_param = sp.ctypes.data_as(ip)
        d.para = _dptr(pa)
        keep = [sp, pa, fi]
        if weights is not None:
            ws = np.ascontiguousarray(np.concatenate([[0], np.cumsum([len(w) for w in weights])]), dtype=np.int32)
            wp = np.ascontiguousarray([p for w in weights for p, _ in w], dtype=np.int32)
            wv = np.ascontiguousarray([v for w in weights for _, v in w], dtype=np.float64)
            d.filter = 1
            d.weight_start, d.weight_pos, d.weight_val = ws.ctypes.data_as(ip), wp.ctypes.data_as(ip), _dptr(wv)
            keep += [ws, wp, wv]
        return cls(model, None, None, material, incidences, _desc=(d, keep))

    def step(self, step=None, beta=0.0, eta=0.5):
        st = None if step is None else np.ascontiguousarray(step, dtype=np.float64)
        if st is not None and len(st) != self.nfree:
            raise TddaError("step needs one value per free parameter")
        _check(lib().tdda_design_step(self.h, None if st is None else _dptr(st), float(beta), float(eta)))

    def get(self):
        """(Para_origin, Para_filtered, Para)."""
        a, b, c = (np.zeros(self.npara) for _ in range(3))
        _check(lib().tdda_design_get(self.h, _dptr(a), _dptr(b), _dptr(c)))
        return a, b, c

    def close(self):
        if self.h:
            lib().tdda_design_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _request(Nx, Ny, Nz, d, lam, t, material, incidences, E0=None, R=None, sources=None, nback=1.0, grid=None):
    """tdda_request of one structure; returns (request, arrays it points
    into, which must stay alive while it is used)."""
    t = _f(t)
    N = len(t) // 3
    g = grid or (0, 0, 0)
    Rp = None
    keep = [t]
    if R is not None:
        R = np.ascontiguousarray(R, dtype=np.int32)
        Rp = R.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        keep.append(R)
    m = _c(material)
    nrhs = len(incidences)
    nK = _f(np.array([i[0] for i in incidences]).reshape(-1))
    nE = _f(np.array([i[1] for i in incidences]).reshape(-1))
    e0 = None if E0 is None else _f(E0)
    keep += [m, nK, nE, e0]
    q = Request()
    q.op = OpDesc(N, Nx, Ny, Nz, Rp, d, lam, nback, g[0], g[1], g[2], 0)
    q.diel_old = _dptr(t)
    q.nmat = len(m)
    q.material = _dptr(m.view(np.float64))
    q.nrhs = nrhs
    q.nK = _dptr(nK)
    q.nE = _dptr(nE)
    q.E0 = _dptr(e0) if e0 is not None else None
    if sources:
        idx = np.ascontiguousarray(np.concatenate([np.asarray(i, np.int32) for i, _ in sources]), dtype=np.int32)
        val = _c(np.concatenate([np.asarray(v) for _, v in sources]))
        q.nsrc = len(sources[0][0])
        q.src_idx = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        q.src_vals = _dptr(val.view(np.float64))
        keep += [idx, val]
    return q, keep


class Server:
    """Request loop with continuous batching on one device (own worker
    thread and stream).  submit() copies the request and returns a ticket;
    result() returns (P[nrhs, 3N], per-RHS solve info, times)."""

    def __init__(self, device: int = 0, budget_bytes: float = 0.0, max_rhs: int = 0, **solve_kw):
        o = ServerOpts()
        lib().tdda_server_opts_default(ctypes.byref(o))
        o.hbm_budget_bytes = budget_bytes
        o.max_rhs = max_rhs
        for k, v in solve_kw.items():
            setattr(o.solve, k, v)
        self.h = ctypes.c_void_p()
        _check(lib().tdda_server_create(device, ctypes.byref(o), ctypes.byref(self.h)))
        self.shape = {}

    def submit(self, Nx, Ny, Nz, d, lam, t, material, incidences, E0=None, R=None, sources=None, nback=1.0,
               grid=None):
        """t: 3N design values; material: complex list after material_rule;
        incidences: [(nK, nE), ...]; sources: None or [(idx, vals), ...] per
        incidence (sparse right-hand sides, same count each)."""
        q, keep = _request(Nx, Ny, Nz, d, lam, t, material, incidences, E0, R, sources, nback, grid)
        N, nrhs = q.op.N, q.nrhs
        tk = ctypes.c_longlong()
        _check(lib().tdda_server_submit(self.h, ctypes.byref(q), ctypes.byref(tk)))
        del keep
        self.shape[tk.value] = (N, nrhs)
        return tk.value

    def poll(self, ticket):
        rc = lib().tdda_server_poll(self.h, ticket)
        if rc < 0:
            _check(rc)
        return rc

    def wait(self, ticket, timeout=-1.0):
        rc = lib().tdda_server_wait(self.h, ticket, float(timeout))
        if rc < 0:
            _check(rc)
        return rc

    def result(self, ticket, want_P=True):
        N, nrhs = self.shape.pop(ticket)
        P = np.zeros((nrhs, 3 * N), np.complex128) if want_P else None
        info = (SolveInfo * nrhs)()
        tm = RequestTimes()
        _check(lib().tdda_server_result(self.h, ticket, _dptr(P.view(np.float64)) if want_P else None, info,
                                        ctypes.byref(tm)))
        return P, _infos(info), dict(t_submit=tm.t_submit, t_admit=tm.t_admit, t_done=tm.t_done,
                                     footprint_bytes=tm.footprint_bytes)

    def stats(self):
        st = ServerStats()
        _check(lib().tdda_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().tdda_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
