"""Fused grain-ensemble lattice intensity (hgx_lattice_intensity) against the oracle composed the
way run_gisaxs composes it (src/sim/hipgisaxs_main.cpp:925-1100): per grain, FF and SF on the
rotated q, DWBA sum with the Fresnel coefficients, |.|^2 accumulated over grains.
Bars: FP64 relL2 <= 1e-10; FP32C relL2 <= 1e-5."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 12.0)
AI = np.radians(0.15)


def _rz(phi):
    c, s = np.cos(phi), np.sin(phi)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def _rx(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


def _oracle_intensity(O, g, shape, dims, la, lb, lc, rep, rots, w, dw, dn2, fc, center=(0, 0, 0)):
    I = np.zeros(g.npix)
    nb = g.nq // g.npix
    for R, wg in zip(rots, w):
        F = dn2 * O.ff_analytic(shape, dims, g, R)
        S = O.sf_lattice(la, lb, lc, rep, center, g, dw_sigma=dw, rot=R)
        I = O.dwba_intensity(g.npix, nb, fc, S, F, wg, I)
    return I


@pytest.fixture(scope="module")
def dwba_grid(oracle_mod):
    O = oracle_mod
    g = O.qgrid_detector(72, 64, 0.172 * 12, 2000.0, 35.5, 0.0, AI, K0)
    g = O.qz_extended(g, K0, AI, 0j)
    fc = O.fresnel(g, [4.88e-6 * (0.10332 / 0.123984) ** 2], [7.37e-8], [0.0], 0, K0, AI)
    return g, fc


def _dev(z):
    import torch

    return torch.from_numpy(np.stack([z.real, z.imag], -1).ravel().copy()).cuda()


@pytest.mark.parametrize("shape", ["sphere", "cylinder"])
def test_lattice_intensity_matches_oracle(gpu_ctx, oracle_mod, dwba_grid, shape):
    from hipgisaxs_amd.api import lattice_intensity

    O = oracle_mod
    g, fc = dwba_grid
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    if shape == "sphere":
        v, w = O.param_distribution(3.0, 7.0, 21, 2, 5.0, 0.5)
        dims = [(v, w / w.sum())]
        a = 20.0
        la, lb, lc = a * np.array([1.0, 0, 0]), a * np.array([0, 1.0, 0]), a * np.array([0.5, 0.5, 0.5])
        rep = (10, 10, 5)
        rots = [_rz(j * np.pi / 4) for j in range(8)]
        wts = np.ones(8)
    else:
        dims = [([3.5, 4.0, 4.5], [0.25, 0.5, 0.25]), ([11.0, 12.0], [0.5, 0.5])]
        a = 15.0
        la, lb, lc = a * np.array([1.0, 0, 0]), a * np.array([0.5, 0.866, 0]), a * np.array([0.5, 0.2887, 0.8165])
        rep = (20, 20, 1)
        rots = [_rx(0.05) @ _rz(0.3), _rz(1.1)]
        wts = np.array([0.7, 0.3])
    dn2 = 2 * 9.4e-6 + 2j * 3.37e-8
    ref = _oracle_intensity(O, g, shape, dims, la, lb, lc, rep, rots, wts, 0.5, dn2, fc)
    fcd = _dev(fc)
    got64 = lattice_intensity(gpu_ctx, shape, dims, la, lb, lc, rep, np.stack(rots), wts, dw_sigma=0.5, dn2=dn2,
                              precision="fp64", fc=fcd).cpu().numpy()
    assert rel_l2(got64, ref) <= 1e-10
    got32 = lattice_intensity(gpu_ctx, shape, dims, la, lb, lc, rep, np.stack(rots), wts, dw_sigma=0.5, dn2=dn2,
                              precision="fp32c", fc=fcd).cpu().numpy()
    err = rel_l2(got32, ref)
    print(f"{shape}: fp32c relL2 {err:.3e}")
    assert err <= 1e-5
    # Born (fc = NULL) and accumulation
    ref_born = _oracle_intensity(O, g, shape, dims, la, lb, lc, rep, rots, wts, 0.5, dn2, None)
    got = lattice_intensity(gpu_ctx, shape, dims, la, lb, lc, rep, np.stack(rots), wts, dw_sigma=0.5, dn2=dn2,
                            precision="fp64")
    got = lattice_intensity(gpu_ctx, shape, dims, la, lb, lc, rep, np.stack(rots), wts, dw_sigma=0.5, dn2=dn2,
                            precision="fp64", accumulate=True, out=got).cpu().numpy()
    assert rel_l2(got, 2 * ref_born) <= 1e-10


def test_lattice_more_grains_than_one_launch(gpu_ctx, oracle_mod, dwba_grid):
    """20 grains > 16 per launch: the host splits launches and accumulates."""
    from hipgisaxs_amd.api import lattice_intensity

    O = oracle_mod
    g, fc = dwba_grid
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    dims = [([5.0], [1.0])]
    a = 20.0
    la, lb, lc = a * np.array([1.0, 0, 0]), a * np.array([0, 1.0, 0]), a * np.array([0.5, 0.5, 0.5])
    rots = [_rz(j * 2 * np.pi / 20) for j in range(20)]
    ref = _oracle_intensity(O, g, "sphere", dims, la, lb, lc, (4, 4, 2), rots, np.ones(20), 0.0, 1.0, fc)
    got = lattice_intensity(gpu_ctx, "sphere", dims, la, lb, lc, (4, 4, 2), np.stack(rots), None, fc=_dev(fc))
    assert rel_l2(got.cpu().numpy(), ref) <= 1e-10
