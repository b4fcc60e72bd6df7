"""Pins the oracle's numeric form factors (approximate and exact triangle methods) to external
truth: closed forms, slice quadrature of the same solid, and exact invariants.

Truth for a square (truncated) pyramid: F(q) = int_0^H 4 a(z) c(z) sinc(qx a) sinc(qy c) e^{i qz z} dz
with half-widths a(z), c(z) linear in z, integrated by 200-point Gauss-Legendre (error ~1e-15).
"""
import numpy as np
import pytest

from conftest import rel_l2

S = lambda x: np.sinc(x / np.pi)  # sin(x)/x


def _slice_pyramid(qx, qy, qz, L, W, H, top_scale):
    """Frustum with base L x W at z=0 shrinking linearly to top_scale * base at z=H."""
    x, w = np.polynomial.legendre.leggauss(200)
    z = (x + 1) * H / 2
    wz = w * H / 2
    f = 1 - (1 - top_scale) * z / H
    a, c = L / 2 * f, W / 2 * f
    return np.array([np.sum(wz * 4 * a * c * S(qx[i] * a) * S(qy[i] * c) * np.exp(1j * qz[i] * z))
                     for i in range(len(qx))])


def _grid(O, n=60, seed=1, scale=1.5):
    rng = np.random.default_rng(seed)
    qx, qy, qz = rng.uniform(-scale, scale, (3, n))
    return O.Grid(n, 1, qx, qy, qz + 0j), qx, qy, qz


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    return (np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]]) @ np.array([[cb, 0, -sb], [0, 1, 0], [sb, 0, cb]])
            @ np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]]))


def test_pyramid_mesh_is_closed_with_exact_volume(oracle_mod):
    from hipgisaxs_amd import meshes

    t = meshes.pyramid()
    assert t.shape == (2048, 9)
    rec = oracle_mod.tri9_to_rec8(t)
    sn = (rec[:, :1] * rec[:, 1:4]).sum(0)
    assert np.max(np.abs(sn)) < 1e-11  # closed surface: sum_t s_t n_t = 0
    assert abs(meshes.mesh_volume(t) - 4000.0 / 3.0) < 1e-9


def test_q_zero_gives_volume(oracle_mod):
    from hipgisaxs_amd import meshes

    O = oracle_mod
    t = meshes.pyramid()
    g = O.Grid(1, 1, [0.0], [0.0], [0j])
    assert abs(O.ff_approx(O.tri9_to_rec8(t), g)[0] - 4000.0 / 3.0) < 1e-9
    assert abs(O.ff_exact(t, g)[0] - 4000.0 / 3.0) < 1e-9


def test_exact_method_matches_slice_quadrature(oracle_mod):
    """External truth for the exact method: square pyramid and truncated pyramid."""
    from hipgisaxs_amd import meshes

    O = oracle_mod
    g, qx, qy, qz = _grid(O)
    ref = _slice_pyramid(qx, qy, qz, 20.0, 20.0, 10.0, 0.0)
    assert rel_l2(O.ff_exact(meshes.pyramid(side_levels=1, base_nx=2, base_ny=2), g), ref) < 1e-10
    t = meshes.truncated_pyramid(16.0, 12.0, 4.0, 54.73)
    d = 4.0 / np.tan(np.radians(54.73))
    # top widths (16-2d, 12-2d) are not a common scale of the base for L != W: integrate directly
    x, w = np.polynomial.legendre.leggauss(200)
    z = (x + 1) * 2.0
    wz = w * 2.0
    a, c = 8.0 - z * d / 4.0, 6.0 - z * d / 4.0
    ref2 = np.array([np.sum(wz * 4 * a * c * S(qx[i] * a) * S(qy[i] * c) * np.exp(1j * qz[i] * z))
                     for i in range(len(qx))])
    assert rel_l2(O.ff_exact(t, g), ref2) < 1e-10


def test_exact_box_matches_closed_form(oracle_mod):
    from hipgisaxs_amd import meshes

    O = oracle_mod
    g, qx, qy, qz = _grid(O, 80, seed=3, scale=2.0)
    L, W, H = 4.0, 6.0, 8.0
    ana = L * W * H * S(qx * L / 2) * S(qy * W / 2) * S(qz * H / 2) * np.exp(1j * qz * H / 2)
    assert rel_l2(O.ff_exact(meshes.box(L, W, H), g), ana) < 1e-9
    # the reference's exp(-iq.r) convention is the complex conjugate for real q
    assert rel_l2(O.ff_exact(meshes.box(L, W, H), g, sign=-1), np.conj(ana)) < 1e-9


def test_approx_method_converges_quadratically(oracle_mod):
    """The centroid rule is second order: the error falls ~4x per midpoint subdivision."""
    from hipgisaxs_amd import meshes

    O = oracle_mod
    g, qx, qy, qz = _grid(O, 40, seed=5, scale=1.0)
    ref = _slice_pyramid(qx, qy, qz, 20.0, 20.0, 10.0, 0.0)
    errs = []
    for lev in (1, 2, 3, 4):
        t = meshes.pyramid(side_levels=lev, base_nx=2 ** lev, base_ny=2 ** lev)
        errs.append(rel_l2(O.ff_approx(O.tri9_to_rec8(t), g), ref))
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert errs[-1] < 5e-3
    assert all(3.5 < r < 5.0 for r in ratios[1:]), ratios  # asymptotic ratio 4


def test_approx_equals_exact_for_one_triangle_limit(oracle_mod):
    """Small q against a fine mesh: centroid rule and exact method agree to O(h^2)."""
    from hipgisaxs_amd import meshes

    O = oracle_mod
    g, *_ = _grid(O, 30, seed=9, scale=0.2)
    t = meshes.pyramid(side_levels=3, base_nx=8, base_ny=8)
    assert rel_l2(O.ff_approx(O.tri9_to_rec8(t), g), O.ff_exact(t, g)) < 1e-2


@pytest.mark.parametrize("method", ["approx", "exact"])
def test_conjugate_symmetry_rotation_translation(oracle_mod, method):
    from hipgisaxs_amd import meshes

    O = oracle_mod
    t = meshes.pyramid(side_levels=2, base_nx=4, base_ny=4)
    f = (lambda tri, g, R=None: O.ff_approx(O.tri9_to_rec8(tri), g, R)) if method == "approx" else \
        (lambda tri, g, R=None: O.ff_exact(tri, g, R))
    g, qx, qy, qz = _grid(O, 40, seed=11)
    gm = O.Grid(40, 1, -qx, -qy, -qz + 0j)
    # F(-q) = conj F(q) for real q
    assert rel_l2(f(t, gm), np.conj(f(t, g))) < 1e-12
    # rotation: F with rot R equals F of the mesh rotated by R^T
    R = _rot(0.4, -0.3, 1.1)
    tr = (t.reshape(-1, 3, 3) @ R).reshape(-1, 9)  # rows v -> R^T v
    assert rel_l2(f(t, g, R), f(tr, g)) < 1e-12
    # translation by d multiplies by exp(i q.d)
    d = np.array([1.5, -2.0, 0.7])
    tt = (t.reshape(-1, 3, 3) + d).reshape(-1, 9)
    assert rel_l2(f(tt, g), f(t, g) * np.exp(1j * (qx * d[0] + qy * d[1] + qz * d[2]))) < 1e-10


def test_as_written_variant_equals_oracle(oracle_mod):
    """The per-pair-allocation timing variant computes the same values as the oracle."""
    from hipgisaxs_amd import meshes

    O = oracle_mod
    rng = np.random.default_rng(9)
    tri = meshes.convex_polyhedron(rng, np.array([5.0, 4.0, 3.0]), 200)
    rec = O.tri9_to_rec8(tri)
    g = O.qz_extended(O.qgrid_explicit(9, 7, -1.0, 0.0, 1.0, 1.0, 0.003, 50.0), 50.0, 0.003, 1e-6 + 1e-8j)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    a = O.ff_approx(rec, g, R)
    b = O.ff_approx_as_written(rec, g, R)
    assert np.array_equal(a, b)


def test_exact_degenerate_case_threshold(oracle_mod):
    """The exact method's degenerate-case threshold on q.n and q.e: with the reference GPU path's 1e-13
    the fine pyramid mesh's form factor is insensitive to a 1e-13 relative change of the vertices (a
    well-conditioned evaluation); with the reference CPU path's DBL_MIN the same change moves it by
    percent (case decisions flip on the last bit of nearly degenerate projections). The GPU kernel
    defaults to the former."""
    import math

    from hipgisaxs_amd import meshes

    O = oracle_mod
    k0 = 2 * math.pi / (1.23984 / 10.0)
    ai = math.radians(0.2)
    tri = meshes.pyramid()
    g = O.qz_extended(O.qgrid_explicit(256, 256, -2.0, 0.0, 2.0, 2.0, ai, k0), k0, ai)
    rows = np.arange(0, 4 * 256, 32)
    gs = O.Grid(256, len(rows), np.repeat(g.qx[:256][None], len(rows), 0).ravel(),
                np.repeat(g.qy[:256][None], len(rows), 0).ravel(), g.qz.reshape(4 * 256, 256)[rows].ravel(),
                g.alpha_f)
    a, b = O.ff_exact(tri, gs, tiny=1e-13), O.ff_exact(tri * (1 + 1e-13), gs, tiny=1e-13)
    c, d = O.ff_exact(tri, gs), O.ff_exact(tri * (1 + 1e-13), gs)
    assert rel_l2(b, a) < 1e-9
    assert rel_l2(d, c) > 1e-3
