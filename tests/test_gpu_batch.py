"""Batched numeric form factors (hgx_batch_*): shapes with uneven meshes, own grids and rotations in one
launch (direct and hipGraph replay), against the oracle and against single-shape calls."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rz(p):
    c, s = np.cos(p), np.sin(p)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


@pytest.mark.parametrize("qpt,bpu", [(4, 16), (2, 1), (8, 3)])
def test_batch_matches_oracle_and_single_calls(gpu_ctx, oracle_mod, qpt, bpu):
    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Batch, Context, tune_batch

    O = oracle_mod
    rng = np.random.default_rng(11)
    sides = [16, 24, 40]
    grids = {}
    for s in sides:
        c = Context(0)
        c.qgrid_explicit(s, s, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
        grids[s] = c
    shapes = []
    for k in range(5):
        tri = meshes.convex_polyhedron(rng, rng.uniform(3, 12, 3), int(rng.integers(100, 900)))
        shapes.append((tri, sides[k % 3], _rz(0.3 * k)))
    try:
        tune_batch(gpu_ctx, qpt, bpu)
        mids = [gpu_ctx.upload_mesh(t, "exact9") for t, _, _ in shapes]
        for prec, tol in (("fp64", 1e-10), ("fp32c", 1e-5)):
            outs = [torch.full((2 * s * s,), float("nan"), dtype=torch.float64, device="cuda") for _, s, _ in shapes]
            items = [(grids[s], mid, R, o) for (t, s, R), mid, o in zip(shapes, mids, outs)]
            b = Batch(gpu_ctx, items, precision=prec)
            for use_graph in (False, True, True):
                for o in outs:
                    o.fill_(float("nan"))
                b.run(use_graph=use_graph)
                torch.cuda.synchronize()
                for (t, s, R), o in zip(shapes, outs):
                    g = O.qgrid_explicit(s, s, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
                    ref = O.ff_approx(O.tri9_to_rec8(t), g, R)
                    got = gpu_ctx.as_complex(o).cpu().numpy()
                    assert np.all(np.isfinite(got))
                    assert rel_l2(got, ref) <= tol, (prec, use_graph, s)
            b.release()
        for m in mids:
            gpu_ctx.release_mesh(m)
    finally:
        tune_batch(gpu_ctx, 4, 16)
        for c in grids.values():
            c.close()


def test_batch_refuses_a_replaced_grid_or_released_mesh(gpu_ctx):
    """A batch's work tables hold the device addresses of its grids and meshes: after a grid is
    replaced or extended, or a mesh released, hgx_batch_run refuses before any launch."""
    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Batch, HgxError

    rng = np.random.default_rng(3)
    gpu_ctx.qgrid_explicit(16, 16, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
    m0 = gpu_ctx.upload_mesh(meshes.convex_polyhedron(rng, rng.uniform(3, 12, 3), 200), "exact9")
    m1 = gpu_ctx.upload_mesh(meshes.convex_polyhedron(rng, rng.uniform(3, 12, 3), 300), "exact9")
    outs = [torch.empty(2 * 256, dtype=torch.float64, device="cuda") for _ in range(2)]
    try:
        b = Batch(gpu_ctx, [(None, m0, None, outs[0]), (None, m1, None, outs[1])], precision="fp64")
        b.run()
        torch.cuda.synchronize()
        gpu_ctx.qz_extended(1e3, 0.0, 0j)
        for g in (False, True):
            with pytest.raises(HgxError, match="create the batch again"):
                b.run(use_graph=g)
        b.release()
        outs = [torch.empty(2 * 4 * 256, dtype=torch.float64, device="cuda") for _ in range(2)]
        b = Batch(gpu_ctx, [(None, m0, None, outs[0]), (None, m1, None, outs[1])], precision="fp64")
        b.run()
        torch.cuda.synchronize()
        gpu_ctx.release_mesh(m1)
        with pytest.raises(HgxError, match="was released"):
            b.run()
        b.release()
    finally:
        gpu_ctx.release_mesh(m0)


def test_batch_on_reference_grids_takes_the_ewald_kernel(gpu_ctx, oracle_mod):
    """Items on QGrid::create q-space grids (src/model/qgrid.cpp:76-84) with DWBA branches, rotated meshes
    of different sizes and series orders, FP32C: one batch (direct and three hipGraph replays) against
    single hgx_ff_numeric calls on the same grids (the same Ewald-grid kernel: equal to FP32 accumulation
    order) and against the oracle (relL2 <= 1e-5)."""
    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Batch, Context

    O = oracle_mod
    k0, ai = 2 * np.pi / (1.23984 / 10.0), np.radians(0.16)
    rng = np.random.default_rng(21)
    specs = {"a": (60, 40, 0.0, 0.0, 3.0, 2.5), "b": (96, 64, -1.0, 0.0, 1.0, 1.0)}
    grids, hgrids = {}, {}
    for key, args in specs.items():
        c = Context(0)
        c.qgrid_qspace(*args, ai, k0)
        c.qz_extended(k0, ai, 0j)
        grids[key] = c
        hgrids[key] = O.qz_extended(O.qgrid_qspace(*args, ai, k0), k0, ai, 0j)
    shapes = []
    for k in range(5):
        tri = meshes.convex_polyhedron(rng, rng.uniform(3, 12, 3), int(rng.integers(200, 1500)))
        shapes.append((tri, "ab"[k % 2], _rz(0.4 * k)))
    try:
        outs, items, singles = [], [], []
        for tri, key, R in shapes:
            gc = grids[key]
            mid = gpu_ctx.upload_mesh(tri, "exact9")
            gmid = gc.upload_mesh(tri, "exact9")
            path, orders = gc.ff_numeric_path(gmid, "approx", "fp32c")
            assert path == "ewald", (path, orders)
            nq = gc.info()["nq"]
            o = torch.full((2 * nq,), float("nan"), dtype=torch.float64, device="cuda")
            outs.append(o)
            items.append((gc, mid, R, o))
            singles.append(gc.as_complex(gc.ff_numeric(gmid, "approx", "fp32c", rot=R)).cpu().numpy())
        b = Batch(gpu_ctx, items, precision="fp32c")
        for use_graph in (False, True, True, True):
            for o in outs:
                o.fill_(float("nan"))
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            for (tri, key, R), o, single in zip(shapes, outs, singles):
                got = gpu_ctx.as_complex(o).cpu().numpy()
                assert np.all(np.isfinite(got))
                ref = O.ff_approx(O.tri9_to_rec8(tri), hgrids[key], R)
                assert rel_l2(got, ref) <= 1e-5, (use_graph, key)
                assert rel_l2(got, single) <= 1e-6, (use_graph, key)
        b.release()
    finally:
        for c in grids.values():
            c.close()
