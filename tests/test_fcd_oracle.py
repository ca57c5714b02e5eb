"""CPU tests that pin the oracle's FCD tensor (AProductCore.cpp:625-695).
The reference's FCD code includes CUDA headers (unbuildable here) and its
Si / Ci tables are not in its repository; the tables are what SiCi/SiCi.py
writes (scipy.special.sici at x = 0.1, 0.2, ...), so the node values are
pinned by construction and the rest by what the mathematics fixes:
* the interpolation is the reference's (nodes exact, linear between);
* gf is the band-limited scalar Green function
  (2 / (pi r)) PV int_0^Kf k sin(k r) / (k^2 - K^2) dk + i sin(K r) / r,
  checked by quadrature, and gf1, gf2 are its r-derivatives (checked by
  finite differences), so a dropped term or wrong sign in the radial
  functions fails;
* the imaginary part of the filtered tensor equals the point tensor's (the
  imaginary part lives on the k = K shell, inside the band);
* far away (Kf r >> 1) the tensor tends to the point tensor;
* each component keeps its parity under x -> -x, y -> -y."""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.special import sici

from oracle import oracle as O

EXACT = sici


def test_table_interpolation_semantics():
    for x in (0.35, 3.14159, 17.77, 123.456):
        s, c = O.sici_table(x)
        a1, a2 = math.floor(x / 0.1) * 0.1, math.ceil(x / 0.1) * 0.1
        (s1, c1), (s2, c2) = sici(a1), sici(a2)
        w = (x - a1) / (a2 - a1)
        assert abs(s - ((1 - w) * s1 + w * s2)) < 1e-13 and abs(c - ((1 - w) * c1 + w * c2)) < 1e-13
    assert O.sici_table(2.0) == tuple(float(v) for v in sici(2.0))


@pytest.mark.parametrize("r", [3.0, 17.0, 55.0, 240.0])
def test_gf_is_band_limited_green_function(r):
    K, d = O.wavenumber(500.0), 10.0
    Kf = math.pi / d
    pv = integrate.quad(lambda k: k * math.sin(k * r) / (k + K), 0.0, Kf, weight="cauchy", wvar=K, limit=400,
                        epsabs=1e-14, epsrel=1e-12)[0]
    want = 2.0 / (math.pi * r) * pv + 1j * math.sin(K * r) / r
    gf = O.fcd_radial(K, d, r, EXACT)[0]
    assert abs(gf - want) < 1e-9 * abs(want)


@pytest.mark.parametrize("r", [3.0, 17.0, 55.0, 240.0])
def test_gf1_gf2_are_derivatives(r):
    K, d = O.wavenumber(620.0), 8.0
    h = 1e-3 * r
    g = [O.fcd_radial(K, d, r + k * h, EXACT)[0] for k in (-2, -1, 0, 1, 2)]
    d1 = (g[0] - 8 * g[1] + 8 * g[3] - g[4]) / (12 * h)
    d2 = (-g[0] + 16 * g[1] - 30 * g[2] + 16 * g[3] - g[4]) / (12 * h * h)
    _, gf1, gf2, _ = O.fcd_radial(K, d, r, EXACT)
    assert abs(gf1 - d1) < 1e-7 * abs(gf1)
    assert abs(gf2 - d2) < 1e-5 * abs(gf2)


def test_imaginary_part_equals_point_tensor():
    K, d = O.wavenumber(500.0), 10.0
    for (x, y, z) in [(10.0, 0, 0), (13.0, -21.0, 30.0), (250.0, 80.0, -130.0)]:
        f, p = O.fcd_inter(K, d, x, y, z), O.green_textbook(K, x, y, z)
        assert np.allclose(f.imag, p.imag, rtol=1e-11, atol=1e-12 * np.abs(p).max())


def test_fcd_tends_to_point_tensor_far_away():
    K, d = O.wavenumber(500.0), 10.0
    errs = []
    for n in (40, 160, 640):
        x, y, z = 0.6 * n * d, 0.3 * n * d, 0.7 * n * d
        f, p = O.fcd_inter(K, d, x, y, z), O.green_textbook(K, x, y, z)
        errs.append(np.linalg.norm(f - p) / np.linalg.norm(p))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 5e-3


def test_fcd_parity():
    K, d = O.wavenumber(542.0), 8.0
    x, y, z = 13.0, -21.0, 30.0
    f = O.fcd_inter(K, d, x, y, z)
    assert np.allclose(O.fcd_inter(K, d, -x, y, z), np.array([1, -1, -1, 1, 1, 1]) * f, rtol=1e-13, atol=0)
    assert np.allclose(O.fcd_inter(K, d, x, -y, z), np.array([1, -1, 1, 1, -1, 1]) * f, rtol=1e-13, atol=0)
    assert np.allclose(O.fcd_inter(K, d, z, y, x)[[5, 4, 2, 3, 1, 0]], f, rtol=1e-13, atol=0)
