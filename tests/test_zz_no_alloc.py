"""hgx_ff_numeric_noalloc (the allocation-free variant of hgx_ff_numeric, partial sums in a fixed
per-device scratch array) against hgx_ff_numeric on the same inputs, through the C ABI: two calls
in a row, the second with a smaller input, for every kernel family (plain FP64, FP32C on the matrix
cores with and without a rotation, FP32C per pair on a non-product grid, the exact method), and
the refusal of a call whose rotated records exceed the array."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rot(a, b):
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1.0]])
    rx = np.array([[1.0, 0, 0], [0, cb, -sb], [0, sb, cb]])
    return rz @ rx


def _noalloc(ctx, mid, method, prec, rot, tv, flags=0):
    import torch

    from hipgisaxs_amd import _native as N

    nq = ctx.info()["nq"]
    out = torch.full((2 * nq,), float("nan"), dtype=torch.float64, device="cuda")
    R = None if rot is None else np.ascontiguousarray(rot, dtype=np.float64).ravel()
    t = None if tv is None else np.ascontiguousarray(tv, dtype=np.float64).ravel()
    m = {"approx": N.FF_APPROX, "exact": N.FF_EXACT}[method]
    p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[prec]
    rc = ctx.lib.hgx_ff_numeric_noalloc(ctx.h, mid, m, p, flags, None if R is None else R.ctypes.data_as(C.c_void_p),
                                        None if t is None else t.ctypes.data_as(C.c_void_p), C.c_void_p(out.data_ptr()))
    return rc, out


# (grid kind, grid side, triangles, method, precision, rotation, flags)
CASES = [
    ("explicit_dwba", 96, 6000, "approx", "fp64", None, 0),
    ("explicit_dwba", 96, 6000, "approx", "fp32c", None, 0),
    ("explicit_dwba", 96, 6000, "approx", "fp32c", (0.4, 0.3), 0),
    ("explicit_born", 200, 20000, "approx", "fp32c", (0.2, 0.1), 0),
    ("detector", 80, 4000, "approx", "fp32c", (0.7, 0.2), 0),
    ("detector", 80, 4000, "approx", "fp64", (0.7, 0.2), 0),
    ("explicit_born", 64, 1500, "exact", "fp64", (0.3, 0.0), 0),
    # 512^2 DWBA: 16 MiB of partial sums per chunk, so the scratch array holds fewer chunks than the
    # workspace path uses
    ("explicit_dwba", 512, 30000, "approx", "fp32c", (0.2, 0.1), 0),
    ("explicit_dwba", 512, 30000, "approx", "fp64", None, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(str(x) for x in c[:5]) + ("-rot" if c[5] else ""))
def test_noalloc_matches_ff_numeric_twice(gpu_ctx, case):
    import torch

    from hipgisaxs_amd import meshes

    kind, side, ntri, method, prec, ang, flags = case
    k0, ai = 2 * np.pi / 0.123984, np.radians(0.2)
    rng = np.random.default_rng(side + ntri)
    tol = 1e-10 if prec == "fp64" else 1e-5
    # the second call smaller; even sides keep rounding-level |q| (q = 0 up to the last bit, where the
    # direct sum is ill-conditioned in any summation order) off the grids
    for call, (s, nt) in enumerate(((side, ntri), (side // 2, ntri // 3))):
        print(f"noalloc {kind} side {s} triangles {nt} {method} {prec} rot {ang} call {call}", flush=True)
        if kind == "detector":
            gpu_ctx.qgrid_detector(s, s, 0.172, 2000.0, s / 2 - 0.5, 0.0, ai, k0)
        else:
            gpu_ctx.qgrid_explicit(s, s, -2.0, 0.0, 2.0, 2.0, ai, k0)
        if kind != "explicit_born":
            gpu_ctx.qz_extended(k0, ai, 0j)
        tri = meshes.convex_polyhedron(rng, rng.uniform(4, 12, 3), nt)
        mid = gpu_ctx.upload_mesh(tri, "exact9")
        R = None if ang is None else _rot(*ang)
        tv = np.array([1.0, -2.0, 0.5])
        try:
            ref = gpu_ctx.ff_numeric(mid, method, prec, rot=R, transvec=tv, flags=flags)
            rc, got = _noalloc(gpu_ctx, mid, method, prec, R, tv, flags)
            torch.cuda.synchronize()
            assert rc == 0, gpu_ctx.lib.hgx_last_error(gpu_ctx.h)
            a, b = got.cpu().numpy(), ref.cpu().numpy()
            assert np.all(np.isfinite(a))
            e = rel_l2(a, b)
            print(f"  relL2 vs hgx_ff_numeric {e:.3e}", flush=True)
            assert e <= tol
        finally:
            gpu_ctx.release_mesh(mid)


def test_noalloc_refuses_records_beyond_the_scratch(gpu_ctx):
    """FP32C on a product grid with a rotation keeps rotated records in the scratch array: a mesh of
    1.1 million triangles needs more than its 32 MiB and is refused before any launch."""
    k0, ai = 2 * np.pi / 0.123984, np.radians(0.2)
    gpu_ctx.qgrid_explicit(16, 16, -2.0, 0.0, 2.0, 2.0, ai, k0)
    rng = np.random.default_rng(2)
    n = 1_100_000
    base = rng.uniform(-10, 10, (n, 3))
    tri = np.concatenate([base, base + [0.3, 0, 0], base + [0, 0.3, 0.05]], axis=1)
    mid = gpu_ctx.upload_mesh(tri, "exact9")
    try:
        rc, _ = _noalloc(gpu_ctx, mid, "approx", "fp32c", _rot(0.3, 0.2), None)
        assert rc == -1
        assert b"scratch" in gpu_ctx.lib.hgx_last_error(gpu_ctx.h)
        rc, _ = _noalloc(gpu_ctx, mid, "approx", "fp32c", None, None)  # no rotation: no records to hold
        assert rc == 0
    finally:
        gpu_ctx.release_mesh(mid)
