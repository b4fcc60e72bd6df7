"""GPU parity (FP64, relL2 <= 1e-10) of q-grids, analytic form factors, the lattice structure
factor, Fresnel coefficients and the DWBA combination against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 12.0)
AI = np.radians(0.15)


def _gpu(ctx, t):
    return ctx.as_complex(t).cpu().numpy()


def _rot(az, ax=0.0):
    c, s = np.cos(az), np.sin(az)
    cx, sx = np.cos(ax), np.sin(ax)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])


@pytest.fixture(scope="module")
def det_grid(oracle_mod):
    O = oracle_mod
    g = O.qgrid_detector(96, 80, 0.172e6 * 8, 2.0e9, 47.5, 0.0, AI, K0)
    return O.qz_extended(g, K0, AI, 0j)


def _upload(ctx, g):
    ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)


def test_qgrid_builders_match_oracle(gpu_ctx, oracle_mod):
    O = oracle_mod
    cases = [
        ("qspace", (70, 50, -1.0, 0.0, 2.0, 2.5, np.radians(0.16), 50.67)),
        ("explicit", (64, 48, -2.0, 0.0, 2.0, 2.0, np.radians(0.2), 50.67)),
        ("detector", (64, 48, 0.172e6, 2.0e9, 31.5, 0.0, AI, K0)),
    ]
    for kind, args in cases:
        getattr(gpu_ctx, f"qgrid_{kind}")(*args)
        gx, gy, gz, ga = gpu_ctx.qgrid_get()
        ref = getattr(O, f"qgrid_{kind}")(*args)
        for a, b in ((gx, ref.qx), (gy, ref.qy), (gz.real, ref.qz.real), (ga, ref.alpha_f)):
            assert np.max(np.abs(a - b)) <= 1e-13 * max(1.0, np.max(np.abs(b))), kind
        gpu_ctx.qz_extended(50.67, np.radians(0.2), 2 * 4.88e-6 + 2j * 7.37e-8)
        gx, gy, gz, ga = gpu_ctx.qgrid_get()
        ref4 = O.qz_extended(ref, 50.67, np.radians(0.2), 2 * 4.88e-6 + 2j * 7.37e-8)
        assert rel_l2(gz, ref4.qz) <= 1e-14


SHAPES = {
    "sphere": [(np.linspace(3, 7, 21), np.exp(-0.5 * ((np.linspace(3, 7, 21) - 5) / 0.5) ** 2))],
    "cylinder": [([3.0, 4.0, 4.5], [0.2, 0.5, 0.3]), ([10.0, 12.0], [0.6, 0.4])],
    "box": [([4.0, 5.0], [1, 1]), ([6.0], [1.0]), ([8.0, 9.0, 10.0], [0.2, 0.3, 0.5])],
    "cube": [([5.0, 6.0], [0.5, 0.5])],
    "prism3": [([6.0, 7.0], [1, 1]), ([5.0], [1.0])],
    "prism6": [([6.0], [1.0]), ([5.0, 5.5], [1, 1])],
    "pyramid": [([16.0], [1.0]), ([16.0, 17.0], [1, 1]), ([4.0], [1.0]), ([54.73], [1.0])],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_analytic_fp64_matches_oracle(gpu_ctx, oracle_mod, det_grid, shape):
    O = oracle_mod
    _upload(gpu_ctx, det_grid)
    R = _rot(0.3, 0.05)
    tv = np.array([1.0, -2.0, 0.5])
    ref = O.ff_analytic(shape, SHAPES[shape], det_grid, R, tv, tau=0.1, eta=0.2)
    got = _gpu(gpu_ctx, gpu_ctx.ff_analytic(shape, SHAPES[shape], "fp64", rot=R, transvec=tv, tau=0.1, eta=0.2))
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= 1e-10


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_analytic_fp32c_matches_oracle(gpu_ctx, oracle_mod, det_grid, shape):
    """FP32C analytic form factors (real q) against the FP64 oracle: relL2 of F and, per pixel, the
    relative error of |F|^2 (floored at 1e-10 of the peak); the prism and pyramid points near their
    cancelling lines take the FP64 formula."""
    O = oracle_mod
    _upload(gpu_ctx, det_grid)
    R = _rot(0.3, 0.05)
    tv = np.array([1.0, -2.0, 0.5])
    ref = O.ff_analytic(shape, SHAPES[shape], det_grid, R, tv, tau=0.1, eta=0.2)
    got = _gpu(gpu_ctx, gpu_ctx.ff_analytic(shape, SHAPES[shape], "fp32c", rot=R, transvec=tv, tau=0.1, eta=0.2))
    assert np.all(np.isfinite(got))
    e = rel_l2(got, ref)
    I, Ir = np.abs(got) ** 2, np.abs(ref) ** 2
    r = np.abs(I - Ir) / np.maximum(Ir, 1e-10 * Ir.max())
    bright = Ir >= 1e-6 * Ir.max()
    print(f"{shape} fp32c relL2 {e:.2e} max pixel {r.max():.2e} (within 6 decades of the peak "
          f"{r[bright].max():.2e})")
    assert e <= 1e-5
    # an FP32 amplitude error of eps |F_peak| is a relative intensity error of about eps sqrt(I_peak / I):
    # 1e-4 within six decades of the peak, 1e-2 down to ten
    assert r[bright].max() <= 1e-4
    assert r.max() <= 1e-2


@pytest.mark.parametrize("precision,dnl", [("fp64", 0j), ("fp64", 2e-6 + 1e-8j), ("fp32c", 0j)])
def test_truncated_sphere_matches_oracle(gpu_ctx, oracle_mod, precision, dnl):
    """Spheres cut to a height (the reference GPU path's 512-slice cap) with radius and height
    distributions, a rotation and a translation, on real and absorbing-layer branches; FP32C evaluates
    the cap in FP64 and the translation phase in FP32. A height above a diameter is refused."""
    from hipgisaxs_amd import _native as N
    from hipgisaxs_amd.api import HgxError

    O = oracle_mod
    g = O.qz_extended(O.qgrid_qspace(40, 30, -1.5, 0.0, 1.5, 1.5, AI, K0), K0, AI, dnl)
    _upload(gpu_ctx, g)
    dims = [([4.5, 5.0], [0.4, 0.6]), ([6.0, 8.5], [0.7, 0.3])]
    R = _rot(0.4, 0.1)
    tv = np.array([1.0, -2.0, 0.5])
    ref = O.ff_truncated_sphere(dims, g, R, tv)
    got = _gpu(gpu_ctx, gpu_ctx.ff_analytic("sphere", dims, precision, rot=R, transvec=tv,
                                            flags=N.SHAPE_TRUNCATED_SPHERE))
    print(f"truncated sphere {precision} dnl={dnl}: relL2 {rel_l2(got, ref):.2e}")
    # FP32C: the cap in FP64, the translation phase in FP32
    assert rel_l2(got, ref) <= (1e-10 if precision == "fp64" else 1e-6)
    # without the flag the heights are not read: the full sphere
    full = _gpu(gpu_ctx, gpu_ctx.ff_analytic("sphere", dims[:1], "fp64", rot=R, transvec=tv))
    assert rel_l2(full, O.ff_analytic("sphere", dims[:1], g, R, tv)) <= 1e-10
    with pytest.raises(HgxError):
        gpu_ctx.ff_analytic("sphere", [([4.0], [1.0]), ([8.5], [1.0])], "fp64", flags=N.SHAPE_TRUNCATED_SPHERE)


def test_analytic_fp32c_refuses_complex_q(gpu_ctx, oracle_mod):
    from hipgisaxs_amd.api import HgxError

    gpu_ctx.qgrid_qspace(16, 12, -1.0, 0.0, 1.0, 1.0, AI, K0)
    gpu_ctx.qz_extended(K0, AI, 2e-6 + 1e-8j)
    with pytest.raises(HgxError, match="real q"):
        gpu_ctx.ff_analytic("sphere", SHAPES["sphere"], "fp32c")


def test_pyramid_reference_compat_matches_oracle(gpu_ctx, oracle_mod, det_grid):
    _upload(gpu_ctx, det_grid)
    d = SHAPES["pyramid"]
    ref = oracle_mod.ff_analytic("pyramid", d, det_grid, compat=1)
    got = _gpu(gpu_ctx, gpu_ctx.ff_analytic("pyramid", d, "fp64", flags=1))
    assert rel_l2(got, ref) <= 1e-10


@pytest.mark.parametrize("lattice", ["bcc", "hcp"])
@pytest.mark.parametrize("compat", [0, 1])
def test_sf_lattice_matches_oracle(gpu_ctx, oracle_mod, det_grid, lattice, compat):
    O = oracle_mod
    _upload(gpu_ctx, det_grid)
    a = 20.0 if lattice == "bcc" else 15.0
    vecs = {"bcc": ([1, 0, 0], [0, 1, 0], [0.5, 0.5, 0.5]), "hcp": ([1, 0, 0], [0.5, 0.866, 0], [0.5, 0.2887, 0.8165])}
    la, lb, lc = (a * np.array(v, dtype=float) for v in vecs[lattice])
    rep = (10, 10, 5) if lattice == "bcc" else (20, 20, 1)
    R = _rot(np.pi / 4)
    ce = np.array([0.5, 1.0, -2.0])
    ref = O.sf_lattice(la, lb, lc, rep, ce, det_grid, dw_sigma=0.5, flags=compat, rot=R)
    got = _gpu(gpu_ctx, gpu_ctx.sf_lattice(la, lb, lc, rep, ce, dw_sigma=0.5, flags=compat, rot=R))
    assert rel_l2(got, ref) <= 1e-10
    # FP32C (HGX_SF_FP32C = 2): FP64-reduced phases, FP32 Dirichlet kernels
    g32 = _gpu(gpu_ctx, gpu_ctx.sf_lattice(la, lb, lc, rep, ce, dw_sigma=0.5, flags=compat | 2, rot=R))
    e32 = rel_l2(g32, ref)
    print(f"sf {lattice} compat={compat} fp32c relL2 {e32:.2e}")
    assert e32 <= 1e-5


def test_fresnel_matches_oracle(gpu_ctx, oracle_mod, det_grid):
    O = oracle_mod
    _upload(gpu_ctx, det_grid)
    cases = [
        ([4.88e-6], [7.37e-8], [0.0], 0),
        ([3.0e-6, 4.88e-6], [1e-8, 7.37e-8], [25.0, 0.0], 0),
        ([3.0e-6, 6.0e-6, 4.88e-6], [1e-8, 2e-8, 7.37e-8], [25.0, 10.0, 0.0], 1),
    ]
    for d, b, t, order in cases:
        ref = O.fresnel(det_grid, d, b, t, order, K0, AI)
        got = _gpu(gpu_ctx, gpu_ctx.fresnel(d, b, t, order, K0, AI))
        assert rel_l2(got, ref) <= 1e-12


def test_dwba_matches_oracle(gpu_ctx, oracle_mod, det_grid):
    O = oracle_mod
    _upload(gpu_ctx, det_grid)
    import torch

    rng = np.random.default_rng(3)
    nq = det_grid.nq
    f = rng.standard_normal(nq) + 1j * rng.standard_normal(nq)
    s = rng.standard_normal(nq) + 1j * rng.standard_normal(nq)
    fc = rng.standard_normal(nq) + 1j * rng.standard_normal(nq)
    dev = lambda z: torch.from_numpy(np.stack([z.real, z.imag], -1).ravel().copy()).cuda()
    I0 = rng.uniform(0, 1, det_grid.npix)
    ref = O.dwba_intensity(det_grid.npix, 4, fc, s, f, 0.7, I0)
    out = torch.from_numpy(I0.copy()).cuda()
    got = gpu_ctx.dwba_intensity(dev(f), dev(fc), dev(s), weight=0.7, accumulate=True, out=out).cpu().numpy()
    assert rel_l2(got, ref) <= 1e-14
