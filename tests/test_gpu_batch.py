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
        tune_batch(qpt, bpu)
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
        tune_batch(4, 16)
        for c in grids.values():
            c.close()
