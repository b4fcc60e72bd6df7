"""The fused ensemble kernel (hgx_ensemble_intensity) and the image helpers through the C ABI:
mixed analytic unit cells with rotations, locations and per-grain centres against the oracle's
per-grain composition (FF sum, SF, DWBA); FP32C cells; the Gaussian smearing against the oracle's
restatement of the reference's convolution."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def _grid(O, gpu_ctx, absorbing):
    g = O.qgrid_qspace(40, 28, -1.5, 0.05, 1.5, 1.2, AI, K0)
    g = O.qz_extended(g, K0, AI, (2e-6 + 3e-8j) if absorbing else 0j)
    fc = O.fresnel(g, [3e-6, 4.88e-6], [1e-8, 7.37e-8], [15.0, 0.0], 0, K0, AI)
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    return g, fc


def _dev(x):
    import torch

    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = np.stack([x.real, x.imag], -1).ravel()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


ELEMENTS = [
    {"shape": "sphere", "dims": [([3.0, 3.5, 4.0], [0.2, 0.5, 0.3])], "rot": np.eye(3),
     "locations": [[0, 0, 0], [5, 1, 2]], "dn2": -(1.2e-5 + 4e-8j)},
    {"shape": "box", "dims": [([12.0], [1.0]), ([8.0], [1.0]), ([6.0], [1.0])], "rot": _rx(0.2) @ _rz(0.4),
     "locations": [[-4, 3, 0]], "dn2": -(1.6e-5 + 2e-7j)},
    {"shape": "pyramid", "dims": [([10.0], [1.0]), ([9.0], [1.0]), ([3.0], [1.0]), ([60.0], [1.0])],
     "rot": _rz(0.1), "locations": [[2, -2, 1]], "dn2": -(8e-6 + 1e-8j)},
]


def _oracle(O, g, fc, elements, la, lb, lc, rep, rots, centers, weights):
    I = np.zeros(g.npix)
    for R, c, w in zip(rots, centers, weights):
        ff = np.zeros(g.nq, complex)
        for e in elements:
            M = R @ e["rot"]
            for t in e["locations"]:
                ff += e["dn2"] * O.ff_analytic(e["shape"], e["dims"], g, rot=M, transvec=t)
        sf = O.sf_lattice(la, lb, lc, rep, c, g, rot=R)
        I = O.dwba_intensity(g.npix, g.nq // g.npix, fc, sf, ff, w, I)
    return I


@pytest.mark.parametrize("absorbing", [False, True])
def test_mixed_cell_matches_oracle(gpu_ctx, oracle_mod, absorbing):
    from hipgisaxs_amd.api import ensemble_intensity

    O = oracle_mod
    g, fc = _grid(O, gpu_ctx, absorbing)
    la, lb, lc, rep = (25.0, 0, 0), (0, 25.0, 0), (12.5, 12.5, 20.0), (4, 3, 2)
    rng = np.random.default_rng(7)
    rots = [_rz(a) @ _rx(b) for a, b in zip(rng.uniform(0, np.pi, 11), rng.uniform(0, 0.1, 11))]
    centers = rng.uniform(-20, 20, (11, 3))
    weights = rng.uniform(0.5, 1.5, 11)
    got = ensemble_intensity(gpu_ctx, ELEMENTS, la, lb, lc, rep, np.stack(rots), centers=centers, weights=weights,
                             fc=_dev(fc)).cpu().numpy()
    ref = _oracle(O, g, fc, ELEMENTS, la, lb, lc, rep, rots, centers, weights)
    r = rel_l2(got, ref)
    print("absorbing", absorbing, "relL2", r)
    assert r <= 1e-10


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-5)])
@pytest.mark.parametrize("case", ["flat", "tilted", "compat", "skew"])
def test_structure_factor_forms(gpu_ctx, oracle_mod, precision, tol, case):
    """The structure factor of the fused kernel in each of its forms, with Debye-Waller damping and
    3-D grain centres: grains rotated about z only (the in-plane lattice directions and the centre
    phase the same on every DWBA branch), tilted grains (every factor per branch), the reference's
    lb[2]*lb[2] term (src/sf/sf.cpp:212, a constant phase), and a grain matrix that is not a rotation
    (the Debye-Waller factor per grain). Against the oracle's per-grain composition."""
    from hipgisaxs_amd.api import ensemble_intensity

    O = oracle_mod
    g, fc = _grid(O, gpu_ctx, False)
    el = [{"shape": "sphere", "dims": [([3.0, 3.5], [0.5, 0.5])], "rot": np.eye(3),
           "locations": [[0, 0, 0]], "dn2": -1.2e-5}]
    # l_b with a z component, so that the reference's lb[2]*lb[2] differs from lb[2]*mq[2]
    la, lb, lc, rep = (20.0, 0, 0), (0, 20.0, 4.0 if case == "compat" else 0.0), (10.0, 10.0, 10.0), (6, 5, 3)
    rng = np.random.default_rng(11)
    angles = rng.uniform(0, 2 * np.pi, 6)
    if case == "flat":
        rots = [_rz(a) for a in angles]
        centers = np.c_[rng.uniform(-30, 30, (6, 2)), np.zeros(6)]
    else:
        rots = [_rz(a) @ _rx(b) for a, b in zip(angles, rng.uniform(0.02, 0.3, 6))]
        centers = rng.uniform(-30, 30, (6, 3))
    if case == "skew":
        rots[2] = rots[2] @ np.diag([1.0, 1.0, 1.02])
    flags = 1 if case == "compat" else 0
    weights = rng.uniform(0.5, 1.5, 6)
    dw = 0.7
    got = ensemble_intensity(gpu_ctx, el, la, lb, lc, rep, np.stack(rots), centers=centers, weights=weights,
                             dw_sigma=dw, sf_flags=flags, fc=_dev(fc), precision=precision).cpu().numpy()
    I = None
    for R, c, w in zip(rots, centers, weights):
        ff = el[0]["dn2"] * O.ff_analytic("sphere", el[0]["dims"], g, rot=R)
        sf = O.sf_lattice(la, lb, lc, rep, c, g, dw_sigma=dw, flags=flags, rot=R)
        I = O.dwba_intensity(g.npix, g.nq // g.npix, fc, sf, ff, w, I)
    r = rel_l2(got, I)
    print(case, precision, "relL2", r)
    assert r <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-5)])
@pytest.mark.parametrize("shape,dims", [
    ("box", [([30.0, 34.0], [0.6, 0.4]), ([10.0, 11.0, 12.0], [0.2, 0.5, 0.3]), ([40.0, 44.0], [0.5, 0.5])]),
    ("cube", [([9.0], [1.0])]),
])
def test_separable_cells(gpu_ctx, oracle_mod, precision, tol, shape, dims):
    """Box and one-node cube cells under grains rotated about z, with an element rotation about z and
    two locations at one height: the kernel pair that evaluates the in-plane factor once per grain and
    the vertical factor once per branch. Debye-Waller damping, 3-D grain centres, DWBA branches."""
    from hipgisaxs_amd.api import ensemble_intensity

    O = oracle_mod
    g, fc = _grid(O, gpu_ctx, False)
    el = [{"shape": shape, "dims": dims, "rot": _rz(0.3), "locations": [[0, 0, 2.0], [25, 7, 2.0]], "dn2": -1e-5}]
    la, lb, lc, rep = (60.0, 0, 0), (0, 45.0, 0), (0, 0, 50.0), (4, 6, 1)
    rng = np.random.default_rng(5)
    rots = [_rz(a) for a in rng.uniform(0, np.pi, 7)]
    centers = rng.uniform(-40, 40, (7, 3))
    weights = rng.uniform(0.5, 1.5, 7)
    got = ensemble_intensity(gpu_ctx, el, la, lb, lc, rep, np.stack(rots), centers=centers, weights=weights,
                             dw_sigma=0.4, fc=_dev(fc), precision=precision).cpu().numpy()
    I = None
    for R, c, w in zip(rots, centers, weights):
        M = R @ el[0]["rot"]
        ff = 0
        for t in el[0]["locations"]:
            ff = ff + el[0]["dn2"] * O.ff_analytic(shape, dims, g, rot=M, transvec=t)
        sf = O.sf_lattice(la, lb, lc, rep, c, g, dw_sigma=0.4, rot=R)
        I = O.dwba_intensity(g.npix, g.nq // g.npix, fc, sf, ff, w, I)
    r = rel_l2(got, I)
    print(shape, precision, "separable relL2", r)
    assert r <= tol


@pytest.mark.parametrize("shape,dims", [
    ("box", [([30.0, 32.0], [0.5, 0.5]), ([10.0], [1.0]), ([40.0, 44.0, 48.0], [0.2, 0.5, 0.3])]),
    ("cube", [([9.0, 11.0], [0.6, 0.4])]),
    ("sphere", [([4.0, 5.0, 6.0], [0.25, 0.5, 0.25])]),
    ("cylinder", [([4.0], [1.0]), ([12.0, 14.0], [0.5, 0.5])]),
    ("prism3", [([6.0, 7.0], [0.5, 0.5]), ([5.0], [1.0])]),
    ("prism6", [([6.0], [1.0]), ([5.0, 5.5], [0.5, 0.5])]),
    ("pyramid", [([16.0], [1.0]), ([16.0, 17.0], [0.5, 0.5]), ([4.0], [1.0]), ([54.73], [1.0])]),
])
def test_fp32c_cells(gpu_ctx, oracle_mod, shape, dims):
    from hipgisaxs_amd.api import ensemble_intensity

    O = oracle_mod
    g, fc = _grid(O, gpu_ctx, False)
    el = [{"shape": shape, "dims": dims, "rot": _rx(0.05), "locations": [[0, 0, 0], [3, 4, 0]], "dn2": -1e-5}]
    la, lb, lc, rep = (40.0, 0, 0), (20.0, 34.64, 0), (0, 0, 50.0), (5, 5, 1)
    rots = np.stack([_rz(a) for a in np.linspace(0, np.pi, 7)])
    got = ensemble_intensity(gpu_ctx, el, la, lb, lc, rep, rots, fc=_dev(fc), precision="fp32c").cpu().numpy()
    ref = _oracle(O, g, fc, el, la, lb, lc, rep, list(rots), np.zeros((7, 3)), np.ones(7))
    r = rel_l2(got, ref)
    print(shape, "fp32c relL2", r)
    assert r <= 1e-5


def test_fp32c_refuses_what_it_does_not_cover(gpu_ctx, oracle_mod):
    from hipgisaxs_amd.api import HgxError, ensemble_intensity

    O = oracle_mod
    _grid(O, gpu_ctx, True)
    with pytest.raises(HgxError, match="real q"):
        ensemble_intensity(gpu_ctx, ELEMENTS[:1], (20, 0, 0), (0, 20, 0), (0, 0, 20), (2, 2, 2), np.eye(3)[None],
                           precision="fp32c")



def test_fp32c_mixed_cell_matches_oracle(gpu_ctx, oracle_mod):
    """A cell of a sphere, a box and a pyramid with rotations, locations and grain centres in FP32C
    (the shape chosen per element; pyramid points near qx qy = 0 in FP64)."""
    from hipgisaxs_amd.api import ensemble_intensity

    O = oracle_mod
    g, fc = _grid(O, gpu_ctx, False)
    la, lb, lc, rep = (25.0, 0, 0), (0, 25.0, 0), (12.5, 12.5, 20.0), (4, 3, 2)
    rng = np.random.default_rng(9)
    rots = [_rz(a) @ _rx(b) for a, b in zip(rng.uniform(0, np.pi, 5), rng.uniform(0, 0.1, 5))]
    centers = rng.uniform(-20, 20, (5, 3))
    weights = rng.uniform(0.5, 1.5, 5)
    els = [dict(e, dn2=complex(e["dn2"]).real) for e in ELEMENTS]
    got = ensemble_intensity(gpu_ctx, els, la, lb, lc, rep, np.stack(rots), centers=centers, weights=weights,
                             fc=_dev(fc), precision="fp32c").cpu().numpy()
    ref = _oracle(O, g, fc, els, la, lb, lc, rep, rots, centers, weights)
    r = rel_l2(got, ref)
    print("mixed fp32c relL2", r)
    assert r <= 1e-5


@pytest.mark.parametrize("compat", [False, True])
@pytest.mark.parametrize("shape,sigma", [((31, 17), 1.0), ((64, 48), 2.3), ((5, 300), 0.7)])
def test_gaussian_smear(gpu_ctx, shape, sigma, compat):
    import torch

    import oracle.hig as OH
    from hipgisaxs_amd.api import gaussian_smear

    nrow, ncol = shape
    img = np.random.default_rng(nrow * ncol).gamma(2.0, 1.0, (nrow, ncol))
    d = torch.from_numpy(img.ravel().copy()).cuda()
    gaussian_smear(gpu_ctx, d, ncol, nrow, sigma, compat)
    got = d.cpu().numpy().reshape(nrow, ncol)
    ref = OH.smear(img, sigma, compat)
    assert rel_l2(got, ref) <= 1e-14
