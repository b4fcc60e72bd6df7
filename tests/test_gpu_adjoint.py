"""Record adjoint of the approximate numeric form factor (hgx_ff_numeric_adjoint) against the oracle's
record adjoint (itself pinned by finite differences in test_oracle_adjoint.py): plain FP64 and FP32C
(VALU kernel) on the reference's non-product grids (detector, QGrid::create q-space), FP32C on product
grids (matrix cores), with rotations, complex DWBA branches, a q = 0 point, a degenerate triangle and
the caller's triangle order."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)


def _rot(a, b):
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1.0]])
    rx = np.array([[1.0, 0, 0], [0, cb, -sb], [0, sb, cb]])
    return rz @ rx


def _mesh(rng, n):
    from hipgisaxs_amd import meshes

    tri = meshes.convex_polyhedron(rng, rng.uniform(4, 9, 3), n)
    # a degenerate triangle in the middle: dropped at upload, its gradient row must be zero
    bad = tri[3].copy().reshape(3, 3)
    bad[2] = bad[1]
    return np.concatenate([tri[:5], bad.reshape(1, 9), tri[5:]])


def _oracle_grad(O, tri, g, W, R):
    keep = [i for i in range(len(tri)) if i != 5]
    rec = O.tri9_to_rec8(tri[keep])
    full = np.zeros((len(tri), 6))
    full[keep] = O.ff_approx_adjoint(rec, g, W, R)
    return full


def test_adjoint_fp64_detector_grid(oracle_mod):
    import torch

    from hipgisaxs_amd.api import Context

    O = oracle_mod
    rng = np.random.default_rng(5)
    tri = _mesh(rng, 500)
    R = _rot(0.4, -0.3)
    c = Context(0)
    try:
        c.qgrid_detector(40, 30, 0.172, 2000.0, 19.5, 0.0, AI, K0)
        c.qz_extended(K0, AI, 3e-6 + 2e-8j)
        g = O.qz_extended(O.qgrid_detector(40, 30, 0.172, 2000.0, 19.5, 0.0, AI, K0), K0, AI, 3e-6 + 2e-8j)
        assert not c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        W = rng.normal(size=g.nq) + 1j * rng.normal(size=g.nq)
        Wd = torch.from_numpy(np.stack([W.real, W.imag], -1).ravel()).cuda()
        got = c.ff_numeric_adjoint(mid, Wd, "fp64", rot=R).cpu().numpy()
        ref = _oracle_grad(O, tri, g, W, R)
        assert np.all(got[5] == 0.0)
        e = rel_l2(got, ref)
        print(f"fp64 adjoint relL2 {e:.2e}")
        assert e <= 1e-10
    finally:
        c.close()


@pytest.mark.parametrize("grid,dwba,rotated,ntri", [
    ("detector", 3e-6 + 2e-8j, True, 500),   # complex DWBA branches
    ("qspace", 0j, False, 1500),             # real branches, a partial last triangle block
    ("qspace", None, True, 3000),            # Born, one branch
])
def test_adjoint_fp32c_any_grid(oracle_mod, grid, dwba, rotated, ntri):
    import torch

    from hipgisaxs_amd.api import Context

    O = oracle_mod
    rng = np.random.default_rng(ntri)
    tri = _mesh(rng, ntri)
    R = _rot(0.4, -0.3) if rotated else None
    c = Context(0)
    try:
        if grid == "detector":
            args = (40, 30, 0.172, 2000.0, 19.5, 0.0, AI, K0)
            c.qgrid_detector(*args)
            g = O.qgrid_detector(*args)
        else:
            # an even column count: no column at q_y = 0, so no q-point within rounding of q = 0, where
            # 1/|q|^2 is ill-posed in any two evaluations of the grid
            args = (56, 41, -2.0, 0.0, 2.0, 2.0, AI, K0)
            c.qgrid_qspace(*args)
            g = O.qgrid_qspace(*args)
        if dwba is not None:
            c.qz_extended(K0, AI, dwba)
            g = O.qz_extended(g, K0, AI, dwba)
        assert not c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        W = rng.normal(size=g.nq) + 1j * rng.normal(size=g.nq)
        Wd = torch.from_numpy(np.stack([W.real, W.imag], -1).ravel()).cuda()
        ref = _oracle_grad(O, tri, g, W, R)
        got64 = c.ff_numeric_adjoint(mid, Wd, "fp64", rot=R).cpu().numpy()
        got32 = c.ff_numeric_adjoint(mid, Wd, "fp32c", rot=R).cpu().numpy()
        e64, e32 = rel_l2(got64, ref), rel_l2(got32, ref)
        # per triangle: the FP32C row against the FP64 row, relative to the largest row of its half
        # (dL/dA, dL/dc), so that no triangle's gradient is lost under the image-wide norm
        d = np.abs(got32 - got64)
        row = max(np.abs(got64[:, :3]).max(), 1e-300), max(np.abs(got64[:, 3:]).max(), 1e-300)
        worst = max(d[:, :3].max() / row[0], d[:, 3:].max() / row[1])
        print(f"{grid} dwba={dwba} rot={rotated} ntri={ntri}: fp64 {e64:.2e} fp32c {e32:.2e} worst {worst:.2e}")
        assert np.all(got32[5] == 0.0)
        assert e64 <= 1e-10
        assert e32 <= 1e-5
        assert worst <= 1e-5
    finally:
        c.close()


@pytest.mark.parametrize("operands", ["f16x2", "f16x2/4", "fp32"])
@pytest.mark.parametrize("ncol,nrow,dwba,rotated", [
    (40, 23, None, False),
    (37, 33, 3e-6 + 2e-8j, True),
    (130, 70, 0j, True),  # real branches; columns past two 64-column groups, rows past a 16-row step
])
def test_adjoint_fp32c_product_grid(oracle_mod, ncol, nrow, dwba, rotated, operands):
    import torch

    from hipgisaxs_amd.api import Context, tune_adjoint

    O = oracle_mod
    rng = np.random.default_rng(ncol + nrow)
    tri = _mesh(rng, 1500)
    R = _rot(0.7, 0.2) if rotated else None
    c = Context(0)
    try:
        g = O.qgrid_explicit(ncol, nrow, -2.0, 0.0, 2.0, 2.0, AI, K0)
        # an exact q = 0 point (F = V there): zero the qy of column 0 and the qz of the last row
        qy = g.qy.copy()
        qy[np.arange(nrow) * ncol] = 0.0
        qz = g.qz.copy()
        qz[(nrow - 1) * ncol:] = 0.0
        g = O.Grid(ncol, nrow, g.qx, qy, qz, g.alpha_f)
        if dwba is not None:
            g = O.qz_extended(g, K0, AI, dwba)
        c.qgrid_set(ncol, nrow, g.qx, g.qy, g.qz, g.alpha_f)
        assert c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        W = rng.normal(size=g.nq) + 1j * rng.normal(size=g.nq)
        Wd = torch.from_numpy(np.stack([W.real, W.imag], -1).ravel()).cuda()
        ref = _oracle_grad(O, tri, g, W, R)
        got64 = c.ff_numeric_adjoint(mid, Wd, "fp64", rot=R).cpu().numpy()
        tune_adjoint(c, operands.split("/")[0], int(operands.split("/")[1]) if "/" in operands else 2)
        got32 = c.ff_numeric_adjoint(mid, Wd, "fp32c", rot=R).cpu().numpy()
        e64, e32 = rel_l2(got64, ref), rel_l2(got32, ref)
        print(f"{ncol}x{nrow} dwba={dwba} rot={rotated} {operands}: fp64 {e64:.2e} fp32c {e32:.2e}")
        assert np.all(got32[5] == 0.0)
        assert e64 <= 1e-10
        assert e32 <= 1e-5
    finally:
        tune_adjoint(c, "f16x2")
        c.close()
