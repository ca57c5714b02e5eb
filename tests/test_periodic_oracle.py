"""CPU tests that pin the oracle's periodic operator (AProductCore.cpp:251-536,
573-586): a direct sum over the image lattice of the textbook tensor, each
image damped by exp(-(0.01 K r)^4).  The reference's periodic code includes
CUDA headers and is not buildable here, and the reference holds no expected
outputs for it, so the restatement is pinned by what the mathematics fixes:
with no images it is the damped free-space operator of the C restatement;
the image sum keeps each component's parity (what the octant storage of the
spectrum relies on); the operator stays complex symmetric; a planted
one-sided image range breaks the parity check."""
import numpy as np

from oracle import oracle as O


def test_no_images_is_the_damped_free_space_operator():
    K = O.wavenumber(500.0)
    R = O.build_a_bulk(3, 2, 2)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(len(R)) + 1j * rng.standard_normal(len(R))
    y0 = O.aproduct_direct(R, 10.0, K, b)                   # C restatement, undamped
    y1 = O.aproduct_periodic_direct(R, 10.0, K, b, 0, 0, 1e9, 1e9)
    rmax = 10.0 * np.sqrt(2 ** 2 + 1 + 1)
    bound = (O.PERIODIC_GAMMA * K * rmax) ** 4              # 1 - exp(-u) <= u
    assert O.rel_l2(y1, y0) <= 4 * bound + 1e-14
    # damping factor of one image, closed form
    g = O.green_periodic(K, 30.0, 0.0, 0.0, 0, 0, 1.0, 1.0)
    assert np.allclose(g, O.green_textbook(K, 30.0, 0.0, 0.0) * np.exp(-(0.01 * K * 30.0) ** 4), rtol=1e-15)


def test_image_sum_keeps_component_parity():
    K = O.wavenumber(542.0)
    x, y, z = 13.0, -7.0, 20.0
    g = O.green_periodic(K, x, y, z, 2, 3, 55.0, 41.0)
    # xx xy xz yy yz zz: xy odd in x and y, xz odd in x and z, yz odd in y and z
    sx = np.array([1, -1, -1, 1, 1, 1])
    sy = np.array([1, -1, 1, 1, -1, 1])
    sz = np.array([1, 1, -1, 1, -1, 1])
    assert np.allclose(O.green_periodic(K, -x, y, z, 2, 3, 55.0, 41.0), sx * g, rtol=1e-13, atol=0)
    assert np.allclose(O.green_periodic(K, x, -y, z, 2, 3, 55.0, 41.0), sy * g, rtol=1e-13, atol=0)
    assert np.allclose(O.green_periodic(K, x, y, -z, 2, 3, 55.0, 41.0), sz * g, rtol=1e-13, atol=0)


def test_planted_one_sided_images_break_parity():
    K = O.wavenumber(542.0)

    def one_sided(x, y, z):
        acc = np.zeros(6, complex)
        for m in range(0, 3):  # planted mistake: m = 0..2 instead of -2..2
            xi = x + m * 55.0
            r = np.sqrt(xi * xi + y * y + z * z)
            acc += O.green_textbook(K, xi, y, z) * np.exp(-(0.01 * K * r) ** 4)
        return acc
    g, gm = one_sided(13.0, 5.0, 2.0), one_sided(-13.0, 5.0, 2.0)
    assert not np.allclose(gm, np.array([1, -1, -1, 1, 1, 1]) * g, rtol=1e-6)


def test_periodic_operator_is_complex_symmetric():
    K = O.wavenumber(500.0)
    R = O.build_a_bulk(3, 3, 2)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(len(R)) + 1j * rng.standard_normal(len(R))
    c = rng.standard_normal(len(R)) + 1j * rng.standard_normal(len(R))
    Ab = O.aproduct_periodic_direct(R, 10.0, K, b, 2, 1, 30.0, 35.0)
    Ac = O.aproduct_periodic_direct(R, 10.0, K, c, 2, 1, 30.0, 35.0)
    assert abs(c @ Ab - b @ Ac) <= 1e-12 * abs(c @ Ab)
