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


@pytest.mark.parametrize("delta,beta", [(9.4e-6, 3.37e-8), (3.0e-6, 2.5e-7), (4.88e-6, 0.0)])
def test_qz_extended_in_a_layer_is_snell_and_decays(oracle_mod, delta, beta):
    """Particle in a layer of index n = 1 - delta + i beta (dnl = 1 - n^2, formed here from n, not
    from 2 delta + 2 i beta). The branches must hold the in-layer incoming and outgoing vertical
    wavevectors: kzi from b2 - b0, kzf from b0 - b1, with
      * b0 + b3 = 0, b1 + b2 = 0 (the four sign pairs of src/model/qgrid.cpp:121-124);
      * dispersion in the layer (Snell: the in-plane component k0 cos(alpha) is conserved),
        kzi^2 + (k0 cos ai)^2 = k0^2 n^2 and kzf^2 + (k0 cos af)^2 = k0^2 n^2 with
        sin af = qz / k0 - sin ai;
      * the incoming wave travels down (Re kzi < 0) and decays downward in an absorbing layer
        (Im kzi < 0: |e^{i kzi z}| -> 0 as z -> -inf); the outgoing wave keeps the direction of the
        vacuum one, sgn(Re kzf) = sgn(qz - k0 sin ai), and decays along it (Im kzf Re kzf >= 0).
    A dropped conjugate, a sign of dnl or a swapped branch fails one of these."""
    O = oracle_mod
    k0, ai = 50.677, np.radians(0.2)
    g = O.qgrid_explicit(9, 11, -1.0, -0.1, 1.0, 2.0, ai, k0)  # qz rows on both sides of k0 sin ai
    n = 1 - delta + 1j * beta
    dnl = 1 - n * n
    e = O.qz_extended(g, k0, ai, dnl).qz.reshape(4, -1)
    assert np.allclose(e[0] + e[3], 0, atol=1e-12) and np.allclose(e[1] + e[2], 0, atol=1e-12)
    kzi = (e[2] - e[0]) / 2
    kzf = (e[0] - e[1]) / 2
    qz = g.qz.real
    kzf0 = qz - k0 * np.sin(ai)
    cos_af2 = 1 - (qz / k0 - np.sin(ai)) ** 2
    scale = k0 * k0
    assert np.allclose(kzi ** 2 + k0 ** 2 * np.cos(ai) ** 2, k0 ** 2 * n * n, rtol=0, atol=1e-11 * scale)
    assert np.allclose(kzf ** 2 + k0 ** 2 * cos_af2, k0 ** 2 * n * n, rtol=0, atol=1e-11 * scale)
    assert np.all(kzi.real < 0)
    sgn = np.where(kzf0 < 0, -1, 1)
    prop = np.abs(kzf.real) > 1e-9 * k0  # without absorption, exit angles below the layer's critical
    assert np.all(np.sign(kzf.real[prop]) == sgn[prop])  # angle give an evanescent kzf (Re = 0)
    if beta == 0:  # (its Im sign then sits on the square root's branch cut: signed zero decides)
        assert np.any(~prop) and np.any(prop)
    else:
        assert np.all(kzi.imag < 0)
        assert np.all(kzf.imag * kzf.real > 0)
    # the vacuum limit: n = 1 gives back the real branches
    e0 = O.qz_extended(g, k0, ai, 0j).qz.reshape(4, -1)
    assert np.allclose(e0[0], qz, atol=1e-12)


def _stack_amplitudes(k0, alpha, n2s, z_if):
    """Field amplitudes of a layer stack solved from the boundary conditions, independently of any
    recursion. Media j = 0 (vacuum) .. M-1 (substrate) with 1 - n_j^2 = n2s[j]; interface j (between
    media j and j + 1) at depth z_if[j]. In medium j the field is T_j e^{-i k_j z} + R_j e^{+i k_j z}
    with k_j = -k0 sqrt(sin^2 alpha - n2s[j]) (the convention of src/model/multilayer.cpp:80-100);
    E and dE/dz are continuous at each interface, the substrate has no R, and T_0 = 1."""
    M = len(n2s)
    k = -k0 * np.sqrt(np.sin(alpha) ** 2 - np.asarray(n2s, complex))
    # unknowns: R_0, (T_j, R_j) for 0 < j < M-1, T_{M-1}
    idx = {}
    cols = 0
    for j in range(M):
        if j > 0:
            idx[("T", j)] = cols
            cols += 1
        if j < M - 1:
            idx[("R", j)] = cols
            cols += 1
    A = np.zeros((2 * (M - 1), cols), complex)
    b = np.zeros(2 * (M - 1), complex)
    for i, z in enumerate(z_if):
        for side, j in ((1.0, i), (-1.0, i + 1)):
            for kind, ph, dsign in (("T", np.exp(-1j * k[j] * z), -1.0), ("R", np.exp(1j * k[j] * z), 1.0)):
                val_e = side * ph
                val_d = side * dsign * 1j * k[j] * ph
                if (kind, j) in idx:
                    A[2 * i, idx[(kind, j)]] += val_e
                    A[2 * i + 1, idx[(kind, j)]] += val_d
                elif kind == "T" and j == 0:  # T_0 = 1 moves to the right-hand side
                    b[2 * i] -= val_e
                    b[2 * i + 1] -= val_d
    x = np.linalg.solve(A, b)
    T = np.array([1.0 + 0j] + [x[idx[("T", j)]] for j in range(1, M)])
    R = np.array([x[idx[("R", j)]] for j in range(M - 1)] + [0j])
    return T, R


@pytest.mark.parametrize("beta1", [0.0, 4e-8])
@pytest.mark.parametrize("order", [0, 1])
def test_fresnel_film_amplitudes_solve_the_boundary_conditions(oracle_mod, order, beta1):
    """Particle inside a film (order 1) and above it (order 0): T_order, R_order of the oracle equal
    the amplitudes obtained by solving the continuity conditions of E and dE/dz at both interfaces
    (T_0 = 1, no upward wave in the substrate), at angles below and above both critical angles,
    with and without absorption in the film."""
    O = oracle_mod
    k0, t = 50.677, 25.0
    d1, d2, b2 = 6.0e-6, 4.88e-6, 7.37e-8
    ai = np.radians(0.23)
    af = np.radians(np.array([0.05, 0.12, 0.18, 0.21, 0.35, 0.6]))
    if beta1 == 0:
        # without absorption, below the film's critical angle its k is imaginary and sits on the
        # square root's branch cut, where the names T and R of the two film waves swap with the
        # sign of a zero imaginary part; the physics is the same, the labels are not
        af = af[af > np.sqrt(2 * d1) * 1.01]
    g = O.Grid(1, af.size, np.zeros(af.size), np.zeros(af.size), np.zeros(af.size) + 0j, af)
    fc = O.fresnel(g, [d1, d2], [beta1, b2], [t, 0.0], order, k0, ai).reshape(4, -1)
    n2s = [0.0, 2 * d1 + 2j * beta1, 2 * d2 + 2j * b2]
    z_if = [0.0, -t]
    Ti, Ri = (v[order] for v in _stack_amplitudes(k0, ai, n2s, z_if))
    for c, a in enumerate(af):
        Tf, Rf = (v[order] for v in _stack_amplitudes(k0, a, n2s, z_if))
        ref = [Ti * np.conj(Tf), Ti * np.conj(Rf), Ri * np.conj(Tf), Ri * np.conj(Rf)]
        assert np.allclose(fc[:, c], ref, rtol=1e-10, atol=1e-13), (order, c)


@pytest.mark.parametrize("order", [1, 2])
def test_fresnel_two_films_without_absorption(oracle_mod, order):
    """Two films on a substrate, real wavevectors in every medium (no absorption, angles above the
    critical angles): the recursion's amplitudes at orders 1 and 2 equal the boundary-condition
    solution. (With absorption in a film that is not the last one the reference's recursion takes
    conj(e^{-i(k_j + k_{j+1}) z}) for e^{+i(k_j + k_{j+1}) z} (src/model/multilayer.cpp:92-95); the two
    agree only for real k, which is the case pinned here.)"""
    O = oracle_mod
    k0 = 50.677
    ds, ts = [2.0e-6, 5.0e-6, 3.0e-6], [12.0, 31.0, 0.0]
    ai = np.radians(0.4)
    af = np.radians(np.array([0.3, 0.45, 0.8]))
    g = O.Grid(1, af.size, np.zeros(af.size), np.zeros(af.size), np.zeros(af.size) + 0j, af)
    fc = O.fresnel(g, ds, [0.0] * 3, ts, order, k0, ai).reshape(4, -1)
    n2s = [0.0] + [2 * d + 0j for d in ds]
    z_if = [0.0, -ts[0], -ts[0] - ts[1]]
    Ti, Ri = (v[order] for v in _stack_amplitudes(k0, ai, n2s, z_if))
    for c, a in enumerate(af):
        Tf, Rf = (v[order] for v in _stack_amplitudes(k0, a, n2s, z_if))
        ref = [Ti * np.conj(Tf), Ti * np.conj(Rf), Ri * np.conj(Tf), Ri * np.conj(Rf)]
        assert np.allclose(fc[:, c], ref, rtol=1e-10, atol=1e-13), (order, c)
