"""Pins the oracle's lattice structure factor, Parratt/Fresnel coefficients, q-grids and the DWBA
combination to direct sums, textbook formulas and exact invariants."""
import numpy as np
import pytest

from conftest import rel_l2

BCC = (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0.5, 0.5, 0.5]))
HCP = (np.array([1.0, 0, 0]), np.array([0.5, 0.866, 0]), np.array([0.5, 0.2887, 0.8165]))


def _small_q_grid(O, lat, a, n=40, seed=4):
    """q with |l_k . q| < pi for every lattice vector, so no branch wrapping happens."""
    rng = np.random.default_rng(seed)
    q = rng.uniform(-1, 1, (n, 3))
    lmax = max(np.linalg.norm(v) for v in lat) * a
    q *= (0.95 * np.pi / lmax) / np.linalg.norm(q, axis=1, keepdims=True) * rng.uniform(0.05, 1, (n, 1))
    return O.Grid(n, 1, q[:, 0], q[:, 1], q[:, 2] + 0j), q


@pytest.mark.parametrize("lat", [BCC, HCP])
def test_sf_equals_centred_direct_lattice_sum(oracle_mod, lat):
    O = oracle_mod
    a = 7.0
    la, lb, lc = (a * v for v in lat)
    N = (4, 3, 2)
    g, q = _small_q_grid(O, lat, a)
    ce = np.array([0.3, -0.2, 0.1])
    got = O.sf_lattice(la, lb, lc, N, ce, g)
    ref = np.zeros(len(q), complex)
    for i in range(N[0]):
        for j in range(N[1]):
            for k in range(N[2]):
                # centred cell index n - (N-1)/2 (the Dirichlet kernel's phase convention)
                r = (i - (N[0] - 1) / 2) * la + (j - (N[1] - 1) / 2) * lb + (k - (N[2] - 1) / 2) * lc
                ref += np.exp(1j * q @ r)
    ref *= np.exp(1j * q @ ce)
    assert rel_l2(got, ref) < 1e-12


def test_sf_dirichlet_identity_with_wrapping(oracle_mod):
    """Large q: the principal-branch pow of the reference equals sin(N x'/2)/sin(x'/2) with
    x' = x wrapped into (-pi, pi]."""
    O = oracle_mod
    rng = np.random.default_rng(8)
    n = 60
    q = rng.uniform(-2, 2, (n, 3))
    g = O.Grid(n, 1, q[:, 0], q[:, 1], q[:, 2] + 0j)
    la, lb, lc = (20.0 * v for v in BCC)
    N = (10, 10, 5)
    got = O.sf_lattice(la, lb, lc, N, [0, 0, 0], g)
    wrap = lambda x: x - 2 * np.pi * np.round(x / (2 * np.pi))
    D = lambda x, m: np.sin(m * wrap(x) / 2) / np.sin(wrap(x) / 2)
    ref = D(q @ la, N[0]) * D(q @ lb, N[1]) * D(q @ lc, N[2])
    assert rel_l2(got, ref) < 1e-10


def test_sf_limits_dw_and_compat(oracle_mod):
    O = oracle_mod
    la, lb, lc = (15.0 * v for v in HCP)
    g = O.Grid(1, 1, [0.0], [0.0], [0j])
    assert abs(O.sf_lattice(la, lb, lc, (20, 20, 1), [0, 0, 0], g)[0] - 400.0) < 1e-9  # x -> 0 limit
    rng = np.random.default_rng(1)
    q = rng.uniform(-1, 1, (30, 3))
    g = O.Grid(30, 1, q[:, 0], q[:, 1], q[:, 2] + 0j)
    s0 = O.sf_lattice(la, lb, lc, (5, 5, 3), [0, 0, 0], g)
    s1 = O.sf_lattice(la, lb, lc, (5, 5, 3), [0, 0, 0], g, dw_sigma=0.5)
    assert rel_l2(s1, s0 * np.exp(-0.5 * 0.25 * (q ** 2).sum(1))) < 1e-13
    # the lb[2]*lb[2] typo only matters when lb has a z component
    assert rel_l2(O.sf_lattice(la, lb, lc, (5, 5, 3), [0, 0, 0], g, flags=1), s0) < 1e-13  # hcp b_z = 0
    lb2 = lb + np.array([0, 0, 2.0])
    assert rel_l2(O.sf_lattice(la, lb2, lc, (5, 5, 3), [0, 0, 0], g, flags=1),
                  O.sf_lattice(la, lb2, lc, (5, 5, 3), [0, 0, 0], g)) > 1e-3


def _kz(k0, a, n2):
    return -k0 * np.sqrt(np.sin(a) ** 2 - n2 + 0j)


def test_fresnel_single_interface_textbook(oracle_mod):
    O = oracle_mod
    k0, ai = 50.677, np.radians(0.2)
    af = np.radians(np.array([0.05, 0.1, 0.17, 0.3, -0.1]))
    g = O.Grid(1, af.size, np.zeros(af.size), np.zeros(af.size), np.zeros(af.size) + 0j, af)
    d, b = 4.88e-6, 7.37e-8
    fc = O.fresnel(g, [d], [b], [0.0], 0, k0, ai).reshape(4, -1)
    n2 = 2 * d + 2j * b
    r = lambda a: (_kz(k0, a, 0) - _kz(k0, a, n2)) / (_kz(k0, a, 0) + _kz(k0, a, n2))
    rf = np.where(af > 0, r(af), 0)
    tf = np.where(af > 0, 1.0, 0)
    assert np.allclose(fc[0], tf, atol=1e-14)
    assert np.allclose(fc[1], np.conj(rf), atol=1e-13)
    assert np.allclose(fc[2], r(ai) * tf, atol=1e-13)
    assert np.allclose(fc[3], r(ai) * np.conj(rf), atol=1e-13)


def test_fresnel_total_reflection_and_vacuum(oracle_mod):
    O = oracle_mod
    k0 = 50.677
    d = 4.88e-6
    ac = np.sqrt(2 * d)
    af = np.array([0.3, 0.6, 0.9]) * ac
    g = O.Grid(1, 3, np.zeros(3), np.zeros(3), np.zeros(3) + 0j, af)
    fc = O.fresnel(g, [d], [0.0], [0.0], 0, k0, 0.5 * ac).reshape(4, -1)
    assert np.allclose(np.abs(fc[1]), 1.0, atol=1e-12)  # |r| = 1 below the critical angle, beta = 0
    assert np.allclose(np.abs(fc[2]), 1.0, atol=1e-12)
    fc0 = O.fresnel(g, [0.0], [0.0], [0.0], 0, k0, 0.5 * ac).reshape(4, -1)
    assert np.allclose(fc0[0], 1) and np.allclose(fc0[1:], 0)  # no contrast: T = 1, R = 0


def test_fresnel_slab_is_airy(oracle_mod):
    """Vacuum / film (thickness t) / substrate: r = (r01 + r12 e^{2i kz1 t}) / (1 + r01 r12 e^{2i kz1 t})
    with kz = -k0 sqrt(sin^2 a - (1 - n^2)) as in src/model/multilayer.cpp:80 (Im kz1 > 0 for an
    absorbing film, so the round-trip factor decays)."""
    O = oracle_mod
    k0, t = 50.677, 30.0
    d1, b1, d2, b2 = 3.0e-6, 2e-8, 4.88e-6, 7.37e-8
    af = np.radians(np.array([0.1, 0.15, 0.22, 0.4]))
    g = O.Grid(1, 4, np.zeros(4), np.zeros(4), np.zeros(4) + 0j, af)
    fc = O.fresnel(g, [d1, d2], [b1, b2], [t, 0.0], 0, k0, af[0]).reshape(4, -1)
    k0z, k1z, k2z = _kz(k0, af, 0), _kz(k0, af, 2 * d1 + 2j * b1), _kz(k0, af, 2 * d2 + 2j * b2)
    r01 = (k0z - k1z) / (k0z + k1z)
    r12 = (k1z - k2z) / (k1z + k2z)
    ph = np.exp(2j * k1z * t)
    airy = (r01 + r12 * ph) / (1 + r01 * r12 * ph)
    assert np.allclose(np.conj(fc[1]), airy, rtol=1e-11, atol=1e-13)


def test_qgrids_on_the_ewald_sphere(oracle_mod):
    O = oracle_mod
    k0, ai = 50.677, np.radians(0.16)
    for g in (O.qgrid_qspace(30, 25, 0.0, 0.0, 3.0, 2.5, ai, k0),
              O.qgrid_detector(32, 24, 0.172, 2000.0, 15.5, 0.0, ai, k0)):
        kf = np.stack([g.qx + k0 * np.cos(ai), g.qy, g.qz.real - k0 * np.sin(ai)])
        assert np.allclose(np.linalg.norm(kf, axis=0), k0, rtol=1e-13)
    g = O.qgrid_qspace(30, 25, 0.0, 0.0, 3.0, 2.5, ai, k0)
    # rows reproduce the requested qz points, highest first (src/model/qgrid.cpp:47-53, 75)
    assert np.allclose(g.qz.real.reshape(25, 30)[:, 0], np.linspace(0, 2.5, 25)[::-1], atol=1e-12)
    e = O.qgrid_explicit(8, 6, -2.0, 0.0, 2.0, 2.0, ai, k0)
    assert np.all(e.qx == 0)
    assert np.allclose(e.qy.reshape(6, 8)[0], np.linspace(-2, 2, 8))
    assert np.allclose(e.qz.real.reshape(6, 8)[:, 3], np.linspace(0, 2, 6)[::-1])


def test_qz_extended_branches(oracle_mod):
    O = oracle_mod
    k0, ai = 50.677, np.radians(0.2)
    g = O.qgrid_explicit(5, 7, -1.0, 0.0, 1.0, 2.0, ai, k0)
    e = O.qz_extended(g, k0, ai, 0j).qz.reshape(4, -1)
    qz = g.qz.real
    s = 2 * k0 * np.sin(ai)
    assert np.allclose(e[0], qz, atol=1e-13)
    assert np.allclose(e[1], s - qz, atol=1e-13)
    assert np.allclose(e[2], qz - s, atol=1e-13)
    assert np.allclose(e[3], -qz, atol=1e-13)


def test_dwba_reduces_to_born_and_is_linear(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(6)
    npix = 50
    ff = rng.standard_normal(4 * npix) + 1j * rng.standard_normal(4 * npix)
    sf = rng.standard_normal(4 * npix) + 1j * rng.standard_normal(4 * npix)
    fc = np.zeros(4 * npix, complex)
    fc[:npix] = 1.0  # delta = beta = 0: T = 1, R = 0
    I = O.dwba_intensity(npix, 4, fc, sf, ff, 0.5)
    assert np.allclose(I, np.abs(0.5 * sf[:npix] * ff[:npix]) ** 2, rtol=1e-14)
    fc2 = rng.standard_normal(4 * npix) + 1j * rng.standard_normal(4 * npix)
    amp = lambda f: np.sqrt(O.dwba_intensity(npix, 4, f, sf, ff, 1.0))
    A = (fc2.reshape(4, -1) * sf.reshape(4, -1) * ff.reshape(4, -1)).sum(0)
    assert np.allclose(amp(fc2), np.abs(A), rtol=1e-13)
