"""Product-grid FP32C numeric form factor on the matrix cores (hgx_qgrid_is_product, product_mfma.hip).

The kernel evaluates the approximate-triangle sum as complex matrix products; it must agree with the
oracle to the FP32C bar on grids whose sides are not multiples of the 32 x 32 tile, on meshes smaller
than one 128-triangle block, under rotation, on complex (DWBA, absorbing layer) branches, with FP32
operands (either complex-product form) and split-fp16 operands, chunked or not, and with the per-pair
VALU kernels (HGX_FLAG_NO_MFMA)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1.0]])
    ry = np.array([[cb, 0, sb], [0, 1.0, 0], [-sb, 0, cb]])
    rx = np.array([[1.0, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return rz @ ry @ rx


def _grid(ctx, O, ncol, nrow, dwba=None):
    ctx.qgrid_explicit(ncol, nrow, -2.0, 0.0, 2.0, 2.0, AI, K0)
    g = O.qgrid_explicit(ncol, nrow, -2.0, 0.0, 2.0, 2.0, AI, K0)
    if dwba is not None:
        ctx.qz_extended(K0, AI, dwba)
        g = O.qz_extended(g, K0, AI, dwba)
    return g


def test_product_grid_detection(oracle_mod):
    from hipgisaxs_amd.api import Context

    O = oracle_mod
    c = Context(0)
    try:
        c.qgrid_explicit(40, 23, -2.0, 0.0, 2.0, 2.0, AI, K0)
        assert c.is_product()
        c.qz_extended(K0, AI, 2e-6 + 1e-8j)
        assert c.is_product()
        c.qgrid_detector(64, 48, 0.172, 2000.0, 31.5, 0.0, AI, K0)
        assert not c.is_product()
        g = O.qgrid_explicit(16, 8, -2.0, 0.0, 2.0, 2.0, AI, K0)
        qy = g.qy.copy()
        c.qgrid_set(16, 8, g.qx, qy, g.qz, g.alpha_f)
        assert c.is_product()
        qy[37] += 1e-9  # one pixel off the product
        c.qgrid_set(16, 8, g.qx, qy, g.qz, g.alpha_f)
        assert not c.is_product()
    finally:
        c.close()


PRODUCT_VARIANTS = [(16, 3, "fp32"), (16, 4, "fp32"), (16, 4, "f16x2", 1, True, False),
                    (16, 4, "f16x2", 2, True, False), (16, 4, "f16x2", 4, True, False),
                    (16, 4, "f16x2", 4, False, False), (16, 4, "f16x2", 2, True, True),
                    (16, 4, "f16x2", 4, False, True), (16, 4, "f16x2", 4, True, True)]


@pytest.mark.parametrize("ncol,nrow,ntri,dwba,rotated", [
    (40, 23, 900, None, False),
    (64, 64, 3000, None, True),
    (33, 31, 60, None, True),           # fewer triangles than one block
    (48, 20, 1500, 2e-6 + 1e-8j, True),  # complex branches (absorbing layer), nz = 4 * 20
    (37, 45, 700, 0j, False),            # real branches
])
@pytest.mark.parametrize("variant", PRODUCT_VARIANTS, ids=lambda v: "x".join(map(str, v)))
def test_product_fp32c_matches_oracle(oracle_mod, ncol, nrow, ntri, dwba, rotated, variant):
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import PRODUCT_DEFAULT, Context, tune_product

    O = oracle_mod
    rng = np.random.default_rng(ncol * 1000 + nrow)
    tri = meshes.convex_polyhedron(rng, rng.uniform(4, 12, 3), ntri)
    R = _rot(0.3, -0.7, 1.1) if rotated else None
    c = Context(0)
    try:
        tune_product(c, *variant)
        g = _grid(c, O, ncol, nrow, dwba)
        assert c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        ref = O.ff_approx(O.tri9_to_rec8(tri), g, R)
        got = c.as_complex(c.ff_numeric(mid, "approx", "fp32c", rot=R)).cpu().numpy()
        valu = c.as_complex(c.ff_numeric(mid, "approx", "fp32c", rot=R, flags=4)).cpu().numpy()
        # The oracle's direct FP64 sum loses about eps / (|q| R) relative accuracy to cancellation:
        # at the rounding-level |q| ~ 1e-16 that odd-sized grids produce, it is not a reference.
        # There F(q) -> V (the mesh volume), which the FP32C small-|q| series gives.
        info = c.mesh_info(mid)
        qn = np.sqrt((np.abs(g.q_rotated(R)) ** 2).sum(0))
        tiny = qn * info["radius"] < 1e-6
        if tiny.any():
            for f in (got, valu):
                assert np.max(np.abs(f[tiny] - info["volume"])) <= 1e-6 * info["volume"]
        ok = ~tiny
        got, valu, ref = got[ok], valu[ok], ref[ok]
        e_mfma, e_valu = rel_l2(got, ref), rel_l2(valu, ref)
        print(f"{ncol}x{nrow} ntri={ntri} dwba={dwba} rot={rotated} {variant}: "
              f"mfma {e_mfma:.2e} valu {e_valu:.2e}")
        assert np.all(np.isfinite(got))
        assert e_mfma <= 1e-5
        assert e_valu <= 1e-5
        c.release_mesh(mid)
    finally:
        tune_product(c, *PRODUCT_DEFAULT)
        c.close()


@pytest.mark.parametrize("operands", ["fp32", "f16x2"])
def test_product_chunked_matches_single_chunk(oracle_mod, operands):
    """Many triangle chunks (partials reduced by the finishing kernel) vs one chunk."""
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import PRODUCT_DEFAULT, Context, tune_numeric, tune_product

    O = oracle_mod
    rng = np.random.default_rng(3)
    tri = meshes.convex_polyhedron(rng, np.array([9.0, 7.0, 5.0]), 6000)
    c = Context(0)
    try:
        g = _grid(c, O, 96, 64)
        mid = c.upload_mesh(tri, "exact9")
        tune_product(c, 16, 4 if operands == "f16x2" else 3, operands)
        tune_numeric(c, 2, 4, 1)  # few waves wanted: one chunk
        one = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy()
        tune_numeric(c, 2, 4, 1 << 16)  # many waves wanted: one block per chunk
        many = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy()
        ref = O.ff_approx(O.tri9_to_rec8(tri), g)
        e1, em, d = rel_l2(one, ref), rel_l2(many, ref), rel_l2(one, many)
        print(f"16-block chunks {e1:.2e}, 1-block chunks {em:.2e}, between {d:.2e}")
        assert e1 <= 1e-5 and em <= 1e-5
        # the two differ only in how many blocks one FP32 accumulation spans
        assert d <= 5e-6
    finally:
        tune_numeric(c, 2, 4, 2048)
        tune_product(c, *PRODUCT_DEFAULT)
        c.close()


@pytest.mark.parametrize("operands", ["fp32", "f16x2"])
def test_product_batch_mixed_with_valu_items(oracle_mod, operands):
    """A batch holding product-grid items (matrix cores) and detector-grid items (per-pair kernels)."""
    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import PRODUCT_DEFAULT, Batch, Context, tune_product

    O = oracle_mod
    rng = np.random.default_rng(21)
    main = Context(0)
    gp = Context(0)
    gd = Context(0)
    try:
        gp.qgrid_explicit(50, 34, -2.0, 0.0, 2.0, 2.0, AI, K0)
        gd.qgrid_detector(40, 30, 0.172, 2000.0, 19.5, 0.0, AI, K0)
        ref_p = O.qgrid_explicit(50, 34, -2.0, 0.0, 2.0, 2.0, AI, K0)
        ref_d = O.qgrid_detector(40, 30, 0.172, 2000.0, 19.5, 0.0, AI, K0)
        shapes = []
        for k in range(4):
            tri = meshes.convex_polyhedron(rng, rng.uniform(3, 10, 3), int(rng.integers(200, 2500)))
            R = _rot(0.2 * k, 0.1 * k, -0.3 * k)
            grid, ref = (gp, ref_p) if k % 2 == 0 else (gd, ref_d)
            shapes.append((tri, R, grid, ref))
        mids = [main.upload_mesh(t, "exact9") for t, _, _, _ in shapes]
        tune_product(main, 16, 4 if operands == "f16x2" else 3, operands)  # read when the batch is made
        outs = [torch.full((2 * gr.info()["nq"],), float("nan"), dtype=torch.float64, device="cuda")
                for _, _, gr, _ in shapes]
        b = Batch(main, [(gr, m, R, o) for (t, R, gr, _), m, o in zip(shapes, mids, outs)], precision="fp32c")
        for use_graph in (False, True, True):
            for o in outs:
                o.fill_(float("nan"))
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            for (t, R, gr, ref), o in zip(shapes, outs):
                want = O.ff_approx(O.tri9_to_rec8(t), ref, R)
                assert rel_l2(main.as_complex(o).cpu().numpy(), want) <= 1e-5, use_graph
        b.release()
    finally:
        tune_product(main, *PRODUCT_DEFAULT)
        for c in (main, gp, gd):
            c.close()


@pytest.mark.parametrize("case", ["mixed_sizes", "far_absorbing", "near_origin"])
def test_product_f16x2_scaling_edges(oracle_mod, case):
    """Split-fp16 operands are scaled by powers of two chosen from |q|, the largest triangle and the
    mesh extent: triangles 1000x smaller than the largest, a mesh 60 nm from the origin on
    absorbing (complex qz) branches, and a grid close to q = 0 must all stay at the FP32C bar."""
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import PRODUCT_DEFAULT, Context, tune_product

    O = oracle_mod
    rng = np.random.default_rng(77)
    tri = meshes.convex_polyhedron(rng, np.array([8.0, 6.0, 5.0]), 400)
    qy0, qz0, qy1, qz1, dwba = -2.0, 0.0, 2.0, 2.0, None
    if case == "mixed_sizes":
        tri = np.concatenate([meshes.subdivide(tri[:4], 5), tri[4:]])
    elif case == "far_absorbing":
        tri = tri + np.tile(np.array([10.0, -20.0, 60.0]), 3)
        dwba = 2e-6 + 2e-7j
    else:
        qy0, qz0, qy1, qz1 = -0.1, 0.0, 0.1, 0.1
    c = Context(0)
    try:
        c.qgrid_explicit(45, 38, qy0, qz0, qy1, qz1, AI, K0)
        g = O.qgrid_explicit(45, 38, qy0, qz0, qy1, qz1, AI, K0)
        if dwba is not None:
            c.qz_extended(K0, AI, dwba)
            g = O.qz_extended(g, K0, AI, dwba)
        assert c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        ref = O.ff_approx(O.tri9_to_rec8(tri), g)
        info = c.mesh_info(mid)
        qn = np.sqrt((np.abs(g.q_rotated(None)) ** 2).sum(0))
        ok = qn * info["radius"] >= 1e-6
        errs = {}
        for variant in PRODUCT_VARIANTS[1:]:
            tune_product(c, *variant)
            got = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy()
            assert np.all(np.isfinite(got))
            errs["x".join(map(str, variant))] = rel_l2(got[ok], ref[ok])
        print(case, {k: f"{v:.2e}" for k, v in errs.items()})
        assert max(errs.values()) <= 1e-5
    finally:
        tune_product(c, *PRODUCT_DEFAULT)
        c.close()


@pytest.mark.parametrize("ncol,nrow,ntri,dwba,rotated,chunks", [
    (40, 23, 900, None, False, False),
    (64, 64, 3000, None, True, True),
    (33, 31, 60, None, True, False),           # fewer triangles than one step or block
    (48, 20, 1500, 2e-6 + 1e-8j, True, True),  # complex branches (absorbing layer), nz = 4 * 20
    (37, 45, 700, 0j, False, False),           # real branches
    (130, 70, 20000, None, True, True),        # several 64 x 64 tiles each way, long chunks
])
@pytest.mark.parametrize("mults", [4, 3])
def test_product_fp64_matches_oracle(oracle_mod, ncol, nrow, ntri, dwba, rotated, chunks, mults):
    """The plain FP64 product-grid kernel (v_mfma_f64_16x16x4_f64, four or three real products per
    complex product) against the oracle at the FP64 bar, and against the per-pair FP64 kernel."""
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Context, tune_numeric, tune_product, tune_product64

    O = oracle_mod
    rng = np.random.default_rng(ncol * 1000 + nrow + 7)
    tri = meshes.convex_polyhedron(rng, rng.uniform(4, 12, 3), ntri)
    R = _rot(0.3, -0.7, 1.1) if rotated else None
    c = Context(0)
    try:
        tune_product64(c, mults)
        if chunks:  # one block per chunk: many partials reduced by the finishing kernel
            tune_product(c, 1, 4, "f16x2", 2, True, True)
            tune_numeric(c, 2, 4, 1 << 16)
        g = _grid(c, O, ncol, nrow, dwba)
        assert c.is_product()
        mid = c.upload_mesh(tri, "exact9")
        ref = O.ff_approx(O.tri9_to_rec8(tri), g, R)
        got = c.as_complex(c.ff_numeric(mid, "approx", "fp64", rot=R)).cpu().numpy()
        pair = c.as_complex(c.ff_numeric(mid, "approx", "fp64", rot=R, flags=4)).cpu().numpy()
        # rounding-level |q| (odd-sized grids): the direct sum is ill-conditioned in both codes
        info = c.mesh_info(mid)
        qn = np.sqrt((np.abs(g.q_rotated(R)) ** 2).sum(0))
        ok = qn * info["radius"] >= 1e-6
        e, e_pair = rel_l2(got[ok], ref[ok]), rel_l2(got[ok], pair[ok])
        print(f"fp64 product {ncol}x{nrow} ntri={ntri} dwba={dwba} rot={rotated} mults={mults}: "
              f"vs oracle {e:.2e}, vs per-pair {e_pair:.2e}")
        assert np.all(np.isfinite(got[ok]))
        assert e <= 1e-10 and e_pair <= 1e-10
        c.release_mesh(mid)
    finally:
        c.close()


def test_product_fp64_batch_matches_single_calls(oracle_mod):
    """FP64 batch items on product grids run the FP64 product kernel: equal to single calls and to
    the oracle, with rotations, on grids of several contexts, graph replay included."""
    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Batch, Context

    O = oracle_mod
    rng = np.random.default_rng(19)
    main = Context(0)
    grids = {s: Context(0) for s in (24, 70)}
    try:
        for s, gc in grids.items():
            gc.qgrid_explicit(s, s, -2.0, 0.0, 2.0, 2.0, AI, K0)
        shapes = [(meshes.convex_polyhedron(rng, rng.uniform(3, 12, 3), int(n)), s, _rot(0.2 * k, 0.1, -0.3 * k))
                  for k, (n, s) in enumerate(((400, 24), (2500, 70), (9000, 70), (150, 24)))]
        mids = [main.upload_mesh(t, "exact9") for t, _, _ in shapes]
        outs = [torch.empty(2 * s * s, dtype=torch.float64, device="cuda") for _, s, _ in shapes]
        b = Batch(main, [(grids[s], m, R, o) for (t, s, R), m, o in zip(shapes, mids, outs)], precision="fp64")
        for use_graph in (False, True, True):
            for o in outs:
                o.fill_(float("nan"))
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            for (t, s, R), o in zip(shapes, outs):
                g = O.qgrid_explicit(s, s, -2.0, 0.0, 2.0, 2.0, AI, K0)
                ref = O.ff_approx(O.tri9_to_rec8(t), g, R)
                got = main.as_complex(o).cpu().numpy()
                assert rel_l2(got, ref) <= 1e-10, (s, use_graph)
        b.release()
    finally:
        main.close()
        for gc in grids.values():
            gc.close()
