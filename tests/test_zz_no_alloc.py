"""GPU test of the allocation-free entry point tdda_aproduct_noalloc: through
the C ABI, its output equals tdda_aproduct's (the entry point it is a
variant of, same passes and tiles) bit for bit, on two calls in a row, the
second with a smaller input, in complex128 and complex64, with and without
the diagonal term; sizes beyond the static scratch array are refused."""
import ctypes

import numpy as np
import pytest

import topodda as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime the library already loaded (the caller's allocator)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevBuf:
    """Device copy of a complex128 host array, allocated by the caller."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        self.n = a.size
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(16, a.nbytes)) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.n, np.complex128)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc(core, b, y, nrhs, al=None, al_per_rhs=0, precision=T.C128):
    L = T.lib()
    L.tdda_aproduct_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return L.tdda_aproduct_noalloc(core.h, b.p, y.p, nrhs, al.p if al is not None else None, al_per_rhs, precision)


def problem(Nx, Ny, Nz, nrhs, seed, sparse=False):
    rng = np.random.default_rng(seed)
    R = O.build_a_bulk(Nx, Ny, Nz)
    if sparse:
        keep = rng.random(len(R) // 3) < 0.6
        R = R.reshape(-1, 3)[keep].reshape(-1)
    n3 = len(R)
    b = rng.standard_normal(nrhs * n3) + 1j * rng.standard_normal(nrhs * n3)
    al = rng.standard_normal(nrhs * n3) + 1j * rng.standard_normal(nrhs * n3)
    return R, b, al


@pytest.mark.parametrize("precision", [T.C128, T.C64])
def test_noalloc_equals_aproduct_two_calls(ctx, precision):
    d, lam = 10.0, 500.0
    # second call: a smaller structure and fewer right-hand sides
    cases = [((20, 18, 10), 3, 11, False), ((12, 8, 6), 1, 12, True)]
    for (Nx, Ny, Nz), nrhs, seed, sparse in cases:
        print(f"noalloc {Nx}x{Ny}x{Nz} nrhs={nrhs} prec={precision} sparse={sparse}", flush=True)
        R, b, al = problem(Nx, Ny, Nz, nrhs, seed, sparse)
        n3 = len(R)
        core = T.AProductCore(ctx, None if not sparse else R, Nx, Ny, Nz, d, lam)
        for diag, per_rhs in ((False, 0), (True, 1), (True, 0)):
            al_use = al if per_rhs else al[:n3]
            ref = core.Aproduct(b.reshape(nrhs, n3) if nrhs > 1 else b,
                                al=(al_use.reshape(nrhs, n3) if per_rhs and nrhs > 1 else al_use) if diag else None,
                                precision=precision)
            ref = np.asarray(ref).reshape(-1)
            db, dy = DevBuf(b), DevBuf(np.zeros_like(b))
            da = DevBuf(al_use) if diag else None
            rc = noalloc(core, db, dy, nrhs, da, per_rhs, precision)
            assert rc == 0, T.lib().tdda_last_error()
            y = dy.get()
            for buf in (db, dy, da):
                if buf is not None:
                    buf.free()
            assert np.array_equal(y, ref), (diag, per_rhs, np.max(np.abs(y - ref)))
            # and the operator bar against the oracle (first right-hand side)
            want = O.aproduct_direct(R, d, O.wavenumber(lam), b[:n3]) + (al_use[:n3] * b[:n3] if diag else 0)
            tol = 1e-12 if precision == T.C128 else 5e-6
            assert O.rel_l2(y[:n3], want) < tol
        core.close()


def test_noalloc_refuses_sizes_beyond_the_scratch(ctx):
    # 80 x 80 x 16 in complex128 needs about 29.5 MB of intermediates per
    # right-hand side: three exceed the 64 MiB array
    core = T.AProductCore(ctx, None, 80, 80, 16, 10.0, 542.0)
    n3 = 3 * 80 * 80 * 16
    db, dy = DevBuf(np.ones(3 * n3)), DevBuf(np.zeros(3 * n3))
    rc = noalloc(core, db, dy, 3)
    assert rc == -5 and b"scratch" in T.lib().tdda_last_error()
    rc = noalloc(core, db, dy, 65)
    assert rc == -1
    # one right-hand side fits
    assert noalloc(core, db, dy, 1) == 0
    db.free()
    dy.free()
    core.close()
