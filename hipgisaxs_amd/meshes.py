"""Triangle-mesh generators for the numeric form-factor path (host-side input preparation).

All generators return vertex triangles as a float64 array of shape (ntri, 9): v1, v2, v3 in
counter-clockwise order seen from outside (outward normals), the ``triangle_t`` layout of the
reference (include/common/typedefs.hpp:102-106). ``hgx_mesh_upload`` derives the 8-real
approximate records (area, normal, centroid) from them as the reference's OBJ reader does
(src/file/objectshape_reader.cpp:271-300).
"""
from __future__ import annotations

import numpy as np


def subdivide(tri: np.ndarray, levels: int) -> np.ndarray:
    """1 -> 4 midpoint subdivision, keeping orientation."""
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3)
    for _ in range(levels):
        a, b, c = t[:, 0], t[:, 1], t[:, 2]
        ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
        t = np.concatenate([np.stack(x, 1) for x in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])
    return t.reshape(-1, 9)


def grid_rect(x0, x1, y0, y1, z, nx, ny, down=True) -> np.ndarray:
    """A z = const rectangle split into nx*ny cells, two triangles each; normal -z if down."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    out = []
    for i in range(nx):
        for j in range(ny):
            p00 = (xs[i], ys[j], z)
            p10 = (xs[i + 1], ys[j], z)
            p11 = (xs[i + 1], ys[j + 1], z)
            p01 = (xs[i], ys[j + 1], z)
            if down:
                out += [p00 + p01 + p11, p00 + p11 + p10]
            else:
                out += [p00 + p10 + p11, p00 + p11 + p01]
    return np.asarray(out, dtype=np.float64)


def pyramid(base: float = 20.0, height: float = 10.0, side_levels: int = 4, base_nx: int = 32,
            base_ny: int = 16) -> np.ndarray:
    """Square pyramid, base [-b/2, b/2]^2 at z = 0, apex (0, 0, h).

    Defaults give the 2048-triangle mesh of the plain-method workload: each lateral face with
    4 levels of midpoint subdivision (4 x 256) and the base as a 32 x 16 grid split on the
    diagonal (1024).
    """
    h = base / 2
    apex = (0.0, 0.0, height)
    corners = [(h, -h, 0.0), (h, h, 0.0), (-h, h, 0.0), (-h, -h, 0.0)]
    faces = [corners[i] + corners[(i + 1) % 4] + apex for i in range(4)]
    side = subdivide(np.asarray(faces), side_levels)
    bottom = grid_rect(-h, h, -h, h, 0.0, base_nx, base_ny, down=True)
    return np.concatenate([side, bottom])


def box(L: float, W: float, H: float, levels: int = 0) -> np.ndarray:
    """Box [-L/2, L/2] x [-W/2, W/2] x [0, H] (the analytic box's frame), 12 * 4^levels triangles."""
    x0, x1, y0, y1, z0, z1 = -L / 2, L / 2, -W / 2, W / 2, 0.0, H
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    tris = []
    for a, b, c, d in quads:
        tris.append(np.concatenate([v[a], v[b], v[c]]))
        tris.append(np.concatenate([v[a], v[c], v[d]]))
    return subdivide(np.asarray(tris), levels)


def prism_polygon(poly_xy: np.ndarray, height: float, levels: int = 0) -> np.ndarray:
    """Right prism over a convex counter-clockwise polygon (z in [0, height])."""
    P = np.asarray(poly_xy, dtype=np.float64)
    n = len(P)
    c = P.mean(0)
    tris = []
    for i in range(n):
        a, b = P[i], P[(i + 1) % n]
        # bottom (normal -z): centre, b, a ; top (normal +z): centre, a, b
        tris.append([c[0], c[1], 0.0, b[0], b[1], 0.0, a[0], a[1], 0.0])
        tris.append([c[0], c[1], height, a[0], a[1], height, b[0], b[1], height])
        # side quad a-b at z=0 and z=h, outward
        tris.append([a[0], a[1], 0.0, b[0], b[1], 0.0, b[0], b[1], height])
        tris.append([a[0], a[1], 0.0, b[0], b[1], height, a[0], a[1], height])
    return subdivide(np.asarray(tris), levels)


def frustum(bottom_xy: np.ndarray, top_xy: np.ndarray, height: float, levels: int = 0) -> np.ndarray:
    """Convex frustum between a bottom polygon (z = 0) and a top polygon (z = height), both CCW."""
    B = np.asarray(bottom_xy, dtype=np.float64)
    T = np.asarray(top_xy, dtype=np.float64)
    n = len(B)
    cb, ct = B.mean(0), T.mean(0)
    tris = []
    for i in range(n):
        a, b = B[i], B[(i + 1) % n]
        ta, tb = T[i], T[(i + 1) % n]
        tris.append([cb[0], cb[1], 0.0, b[0], b[1], 0.0, a[0], a[1], 0.0])
        tris.append([ct[0], ct[1], height, ta[0], ta[1], height, tb[0], tb[1], height])
        tris.append([a[0], a[1], 0.0, b[0], b[1], 0.0, tb[0], tb[1], height])
        tris.append([a[0], a[1], 0.0, tb[0], tb[1], height, ta[0], ta[1], height])
    return subdivide(np.asarray(tris), levels)


def truncated_pyramid(L: float, W: float, H: float, angle_deg: float, levels: int = 0) -> np.ndarray:
    """Square-based truncated pyramid with base angle (degrees), base centred at the origin."""
    d = H / np.tan(np.radians(angle_deg))
    B = np.array([[-L / 2, -W / 2], [L / 2, -W / 2], [L / 2, W / 2], [-L / 2, W / 2]])
    T = np.array([[-L / 2 + d, -W / 2 + d], [L / 2 - d, -W / 2 + d], [L / 2 - d, W / 2 - d], [-L / 2 + d, W / 2 - d]])
    return frustum(B, T, H, levels)


def orient_outward(tri: np.ndarray, inside_point) -> np.ndarray:
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3).copy()
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    flip = np.einsum("ij,ij->i", n, t.mean(1) - np.asarray(inside_point)) < 0
    t[flip] = t[flip][:, [0, 2, 1]]
    return t.reshape(-1, 9)


def convex_polyhedron(rng: np.random.Generator, semi_axes, target_tri: int) -> np.ndarray:
    """Convex hull of points on an ellipsoid (about target_tri triangles: a hull of n points in
    general position has 2n - 4 faces), by the library's quickhull (hgx_hig_convex_hull)."""
    from hipgisaxs_amd.hig import convex_hull

    n = max(4, int(round(target_tri / 2 + 2)))
    x = rng.standard_normal((n, 3))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    x *= np.asarray(semi_axes, dtype=np.float64)
    tri = x[convex_hull(x)].reshape(-1, 9)
    return orient_outward(tri, x.mean(0))


def c3_shapes(nshapes: int = 64, seed: int = 20261206):
    """The batched-polyhedra workload: per shape, semi-axes U[3, 15]^3 nm, a target triangle count
    log-uniform on [500, 50000], a square grid side from {128, 256, 512, 1024}, and the hull of points
    on that ellipsoid. Returns dicts with tri, side, axes and target."""
    import math

    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(nshapes):
        axes = rng.uniform(3.0, 15.0, 3)
        target = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        side = int(rng.choice([128, 256, 512, 1024]))
        tri = convex_polyhedron(rng, axes, target)
        shapes.append({"tri": tri, "side": side, "axes": axes, "target": target})
    return shapes


def mesh_volume(tri: np.ndarray) -> float:
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3)
    return float(np.einsum("ij,ij->", t[:, 0], np.cross(t[:, 1], t[:, 2])) / 6.0)


def trapezoid_line(top, angle, height, length, n_bottom, n_side, n_top, n_length):
    """Closed mesh of a line of trapezoidal cross-section: base [-wb/2, wb/2] on z = 0 with
    wb = top + 2 height cot(angle), top [-top/2, top/2] at z = height, extruded along x over
    [-length/2, length/2]. Perimeter in n_bottom + 2 n_side + n_top segments (counter-clockwise in
    y-z), n_length segments along x, two triangles per quad, caps fanned from the corner centroid
    (0, height/2). Same triangles as the device mesher of the grating fit. Returns [ntri, 9]."""
    wb = top + 2.0 * height / np.tan(angle)
    C = np.array([[-wb / 2, 0.0], [wb / 2, 0.0], [top / 2, height], [-top / 2, height]])
    P = []
    for a, n in ((0, n_bottom), (1, n_side), (2, n_top), (3, n_side)):
        b = (a + 1) % 4
        for i in range(n):
            u = i / n
            P.append((1 - u) * C[a] + u * C[b])
    P = np.array(P)
    npr = len(P)
    xs = -length / 2 + length * np.arange(n_length + 1) / n_length
    tris = []
    for i in range(n_length):
        for k in range(npr):
            k1 = (k + 1) % npr
            a = (xs[i], *P[k])
            b = (xs[i], *P[k1])
            c = (xs[i + 1], *P[k1])
            d = (xs[i + 1], *P[k])
            tris.append((*a, *b, *c))
            tris.append((*a, *c, *d))
    for ci, x in ((0, -length / 2), (1, length / 2)):
        M = (x, 0.0, height / 2)
        for k in range(npr):
            k1 = (k + 1) % npr
            pa, pb = (P[k], P[k1]) if ci else (P[k1], P[k])
            tris.append((*M, x, *pa, x, *pb))
    return np.array(tris, dtype=np.float64)
