"""GPU parity of the Ewald-sphere matrix-core path against the FP64 oracle, through the C ABI.

The path runs the FP32C approximate-triangle form factor (src/ff/cpu/ff_tri_cpu.cpp:200-246) on
grids of the form q = (a_i + c_i vx_j, c_i vy_j, qz(branch, row i)): QGrid::create's q-space grids
(src/model/qgrid.cpp:76-84), detector grids, and any uploaded grid of that form. Cases: the grids of
examples/01-pyramid.hig and examples/02-spheres-bcc.hig at reduced resolution (DWBA, Born, absorbing
layer), a detector grid, grids built to need 0, 1, 2 or 3 series orders, rotated and off-centre meshes.

Bars: relL2 <= 1e-5 against the oracle (the FP32C tolerance); the largest pixel error <= 1e-5 of the
largest |F|; and no worse than twice the per-pair FP32C kernel's own error on the same case, plus 1e-7.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _k0(kev):
    return 2 * np.pi / (1.23984 / kev)


def _rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


# grid cases: (builder, arguments, incidence angle in degrees, keV, dnl of the particle layer or None
# for the Born grid)
GRIDS = {
    # examples/01-pyramid.hig: q-space [0, 0]-[3, 2.5] 1/nm, 10 keV, alpha_i 0.16 deg, DWBA
    "ex01_dwba": ("qspace", (150, 125, 0.0, 0.0, 3.0, 2.5), 0.16, 10.0, 0j),
    # examples/02-spheres-bcc.hig: q-space [-1, 0]-[1, 1], 7.35 keV, alpha_i 0.175 deg, Born
    "ex02_born": ("qspace", (300, 150, -1.0, 0.0, 1.0, 1.0), 0.175, 7.35, None),
    # the same with an absorbing particle layer: complex qz branches
    "ex02_absorbing": ("qspace", (120, 150, -1.0, 0.0, 1.0, 1.0), 0.175, 7.35, complex(4.9e-6, 7.4e-8)),
    # detector geometry: 172 um pixels at 2 m, 12 keV, alpha_i 0.15 deg
    "detector": ("detector", (128, 128, 172e-6, 2.0, 64.0, 20.0), 0.15, 12.0, 0j),
}


def _build(ctx, O, case):
    kind, args, ai_deg, kev, dnl = GRIDS[case]
    ai, k0 = np.radians(ai_deg), _k0(kev)
    if kind == "qspace":
        ctx.qgrid_qspace(*args, ai, k0)
        g = O.qgrid_qspace(*args, ai, k0)
    else:
        ctx.qgrid_detector(*args, ai, k0)
        g = O.qgrid_detector(*args, ai, k0)
    if dnl is not None:
        ctx.qz_extended(k0, ai, dnl)
        g = O.qz_extended(g, k0, ai, dnl)
    return g


@pytest.fixture(scope="module")
def meshes_up(gpu_ctx):
    from hipgisaxs_amd import meshes

    rng = np.random.default_rng(20261204)
    pyr = meshes.pyramid()
    poly = meshes.convex_polyhedron(rng, (12.0, 9.0, 7.0), 4000)
    shifted = pyr + np.tile([40.0, -25.0, 15.0], 3)
    out = {}
    for name, tri in (("pyramid", pyr), ("polyhedron", poly), ("shifted", shifted)):
        out[name] = (tri, gpu_ctx.upload_mesh(tri, "exact9"))
    yield out
    for _, mid in out.values():
        gpu_ctx.release_mesh(mid)


def _check(ctx, O, g, tri, mid, R, want_path="ewald"):
    from hipgisaxs_amd import _native as N

    path, orders = ctx.ff_numeric_path(mid, "approx", "fp32c")
    assert path == want_path, (path, orders)
    ref = O.ff_approx(O.tri9_to_rec8(tri), g, R)
    got = ctx.as_complex(ctx.ff_numeric(mid, "approx", "fp32c", rot=R)).cpu().numpy()
    pair = ctx.as_complex(ctx.ff_numeric(mid, "approx", "fp32c", rot=R, flags=N.FLAG_NO_MFMA)).cpu().numpy()
    e_got, e_pair = rel_l2(got, ref), rel_l2(pair, ref)
    peak = np.abs(ref).max()
    e_max = float(np.abs(got - ref).max() / peak)
    e_pmax = float(np.abs(pair - ref).max() / peak)
    print(f"path {path} orders {orders}: relL2 {e_got:.3e} (per-pair {e_pair:.3e}), "
          f"max/peak {e_max:.3e} (per-pair {e_pmax:.3e})")
    assert e_got <= 1e-5
    assert e_max <= 1e-5
    assert e_got <= 2.0 * e_pair + 1e-7
    return orders


@pytest.mark.parametrize("case", list(GRIDS))
@pytest.mark.parametrize("mesh", ["pyramid", "polyhedron"])
def test_reference_grids_match_oracle(gpu_ctx, oracle_mod, meshes_up, case, mesh):
    g = _build(gpu_ctx, oracle_mod, case)
    assert not gpu_ctx.is_product()
    ok, dv = gpu_ctx.is_ewald()
    assert ok and dv > 0
    tri, mid = meshes_up[mesh]
    _check(gpu_ctx, oracle_mod, g, tri, mid, None)


@pytest.mark.parametrize("case", ["ex01_dwba", "ex02_absorbing"])
@pytest.mark.parametrize("mesh", ["pyramid", "shifted"])
def test_rotated_and_offcentre_meshes(gpu_ctx, oracle_mod, meshes_up, case, mesh):
    g = _build(gpu_ctx, oracle_mod, case)
    tri, mid = meshes_up[mesh]
    _check(gpu_ctx, oracle_mod, g, tri, mid, _rot(0.3, -0.2, 0.7))


@pytest.mark.parametrize("case", ["ex01_dwba", "ex02_absorbing"])
def test_both_schedules_agree(gpu_ctx, oracle_mod, meshes_up, case):
    """The pipelined schedule (column operands a step ahead) against the default one: both within the
    bar, and equal to each other up to FP32 accumulation order (they run the same products)."""
    g = _build(gpu_ctx, oracle_mod, case)
    tri, mid = meshes_up["polyhedron"]
    outs = []
    try:
        for pipe in (1, 0):
            assert gpu_ctx.lib.hgx_tune_ewald(gpu_ctx.h, pipe) == 0
            _check(gpu_ctx, oracle_mod, g, tri, mid, _rot(0.1, 0.4, -0.3))
            outs.append(gpu_ctx.as_complex(gpu_ctx.ff_numeric(mid, "approx", "fp32c",
                                                              rot=_rot(0.1, 0.4, -0.3))).cpu().numpy())
    finally:
        gpu_ctx.lib.hgx_tune_ewald(gpu_ctx.h, 0)
    assert rel_l2(outs[0], outs[1]) <= 1e-12


def _ewald_grid(ncol, nrow, af, tth, ai, k0):
    """Host arrays of QGrid::create's formula (src/model/qgrid.cpp:80-82) for given exit angles."""
    A, T = np.meshgrid(af, tth, indexing="ij")
    qx = k0 * (np.cos(A) * np.cos(T) - np.cos(ai))
    qy = k0 * (np.cos(A) * np.sin(T))
    qz = k0 * (np.sin(A) + np.sin(ai)) + 0j
    return qx.ravel(), qy.ravel(), qz.ravel()


@pytest.mark.parametrize("span_deg,want", [(0.0, 0), (0.02, 1), (0.6, 2), (1.0, 3), (20.0, -1)])
def test_series_orders_follow_the_row_spread(gpu_ctx, oracle_mod, meshes_up, span_deg, want):
    """Uploaded grids of the reference's form whose exit angles span span_deg above 1 deg: the row
    spread of c = cos(af) sets the series orders (0 for equal rows, per-pair kernels past 3)."""
    O = oracle_mod
    k0, ai = _k0(10.0), np.radians(0.2)
    nrow, ncol = 64, 96
    af = np.radians(1.0 + span_deg * np.linspace(0.0, 1.0, nrow))
    tth = np.radians(np.linspace(-3.0, 3.0, ncol))
    qx, qy, qz = _ewald_grid(ncol, nrow, af, tth, ai, k0)
    gpu_ctx.qgrid_set(ncol, nrow, qx, qy, qz, af)
    g = O.Grid(ncol, nrow, qx, qy, qz, af)
    ok, dv = gpu_ctx.is_ewald()
    assert ok and not gpu_ctx.is_product()
    tri, mid = meshes_up["pyramid"]
    path, orders = gpu_ctx.ff_numeric_path(mid, "approx", "fp32c")
    assert orders == want, (orders, dv)
    _check(gpu_ctx, O, g, tri, mid, None, want_path="ewald" if want >= 0 else "pair")


def test_grid_form_is_detected_not_assumed(gpu_ctx, oracle_mod, meshes_up):
    """A perturbed copy of a reference-form grid is not an Ewald grid, FP64 and the exact method never
    take the path, and the product grid keeps its own kernel."""
    from hipgisaxs_amd import _native as N

    k0, ai = _k0(10.0), np.radians(0.2)
    af = np.radians(np.linspace(0.0, 2.0, 40))
    tth = np.radians(np.linspace(-2.0, 2.0, 50))
    qx, qy, qz = _ewald_grid(50, 40, af, tth, ai, k0)
    tri, mid = meshes_up["pyramid"]
    gpu_ctx.qgrid_set(50, 40, qx, qy, qz, af)
    assert gpu_ctx.is_ewald()[0]
    assert gpu_ctx.ff_numeric_path(mid, "approx", "fp64")[0] == "pair"
    assert gpu_ctx.ff_numeric_path(mid, "exact", "fp32c")[0] == "pair"
    assert gpu_ctx.ff_numeric_path(mid, "approx", "fp32c", flags=N.FLAG_NO_MFMA)[0] == "pair"
    qx2 = qx.copy()
    qx2[7 * 50 + 11] += 1e-6
    gpu_ctx.qgrid_set(50, 40, qx2, qy, qz, af)
    assert not gpu_ctx.is_ewald()[0]
    assert gpu_ctx.ff_numeric_path(mid, "approx", "fp32c")[0] == "pair"
    gpu_ctx.qgrid_explicit(50, 40, -2.0, 0.0, 2.0, 2.0, ai, k0)
    assert gpu_ctx.is_product()
    assert gpu_ctx.ff_numeric_path(mid, "approx", "fp32c")[0] == "product"


def test_noalloc_entry_point_takes_the_same_path(gpu_ctx, oracle_mod, meshes_up):
    import torch

    g = _build(gpu_ctx, oracle_mod, "ex01_dwba")
    tri, mid = meshes_up["polyhedron"]
    assert gpu_ctx.ff_numeric_path(mid, "approx", "fp32c")[0] == "ewald"
    a = gpu_ctx.ff_numeric(mid, "approx", "fp32c")
    b = torch.empty_like(a)
    rc = gpu_ctx.lib.hgx_ff_numeric_noalloc(gpu_ctx.h, mid, 0, 1, 0, None, None, b.data_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    ref = oracle_mod.ff_approx(oracle_mod.tri9_to_rec8(tri), g)
    assert rel_l2(gpu_ctx.as_complex(b).cpu().numpy(), ref) <= 1e-5
