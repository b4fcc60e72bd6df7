"""Forward-mode derivatives through the C ABI: hgx_ensemble_intensity_tangent against central
differences of hgx_ensemble_intensity (every shape, every kind of table tangent, real and complex q),
hgx_image_distance_tangent against central differences of hgx_image_distance (every metric), and the
.hig fit gradient (hgx_hig_fit_gradient) against central differences of the oracle's objective;
LMVM on tangent passes, and the particle swarm with one synchronisation per generation against the
particle-by-particle run."""
import copy

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _drz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[-s, -c, 0.0], [c, -s, 0.0], [0.0, 0.0, 0.0]])


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def _grid(O, ctx, absorbing):
    g = O.qgrid_qspace(36, 24, -1.5, 0.05, 1.5, 1.2, AI, K0)
    g = O.qz_extended(g, K0, AI, (2e-6 + 3e-8j) if absorbing else 0j)
    fc = O.fresnel(g, [3e-6, 4.88e-6], [1e-8, 7.37e-8], [15.0, 0.0], 0, K0, AI)
    ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    return g, fc


def _dev(x):
    import torch

    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = np.stack([x.real, x.imag], -1).ravel()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


CELLS = {
    "sphere_box_pyramid": [
        {"shape": "sphere", "dims": [([3.0, 3.5, 4.0], [0.2, 0.5, 0.3])], "rot": np.eye(3),
         "locations": [[0, 0, 0], [5, 1, 2]], "dn2": -(1.2e-5 + 4e-8j)},
        {"shape": "box", "dims": [([12.0], [1.0]), ([8.0], [1.0]), ([6.0], [1.0])], "rot": _rx(0.2) @ _rz(0.4),
         "locations": [[-4, 3, 0]], "dn2": -(1.6e-5 + 2e-7j)},
        {"shape": "pyramid", "dims": [([10.0], [1.0]), ([9.0], [1.0]), ([3.0], [1.0]), ([60.0], [1.0])],
         "rot": _rz(0.1), "locations": [[2, -2, 1]], "dn2": -(8e-6 + 1e-8j)},
    ],
    "cylinder_cube_prisms": [
        {"shape": "cylinder", "dims": [([4.0, 4.5], [0.6, 0.4]), ([10.0, 12.0], [0.5, 0.5])], "rot": _rx(0.1),
         "locations": [[0, 0, 0]], "dn2": -(1.0e-5 + 1e-8j)},
        {"shape": "cube", "dims": [([6.0], [1.0])], "rot": _rz(0.3), "locations": [[3, 3, 0]], "dn2": -1e-5},
        {"shape": "prism3", "dims": [([7.0], [1.0]), ([5.0], [1.0])], "tau": 0.1, "eta": 0.2,
         "locations": [[-3, 2, 1]], "dn2": -(9e-6 + 2e-8j)},
        {"shape": "prism6", "dims": [([6.0], [1.0]), ([4.0], [1.0])], "tau": 0.05, "eta": -0.3,
         "locations": [[1, -4, 0]], "dn2": -(7e-6 + 1e-8j)},
    ],
}


def _model(cell, rng):
    angles = rng.uniform(0, np.pi, 5)
    tilts = rng.uniform(0, 0.1, 5)
    return {"elements": cell, "la": (25.0, 0, 0), "lb": (0, 25.0, 0), "lc": (12.5, 12.5, 20.0), "repet": (4, 3, 2),
            "rots": np.stack([_rz(a) @ _rx(b) for a, b in zip(angles, tilts)]),
            "centers": rng.uniform(-20, 20, (5, 3)), "weights": rng.uniform(0.5, 1.5, 5), "dw_sigma": 0.4,
            "_angles": angles, "_tilts": tilts}


def _tangents(model, rng):
    """One direction per kind of table: node values, node weights, tau/eta, element rotations,
    locations, contrasts, lattice vectors, grain rotations (d/d angle), centres, weights, dw_sigma."""
    els = model["elements"]
    zero_el = [{} for _ in els]

    def with_elements(f):
        out = []
        for e in els:
            out.append(f(e))
        return {"elements": out}

    T = [
        with_elements(lambda e: {"dims": [(rng.uniform(0.2, 1.0, len(v)), np.zeros(len(w))) for v, w in e["dims"]]}),
        with_elements(lambda e: {"dims": [(np.zeros(len(v)), rng.uniform(-0.3, 0.3, len(w))) for v, w in e["dims"]]}),
        with_elements(lambda e: {"tau": 0.7, "eta": -0.4, "rot": _drz(0.37) @ np.asarray(e.get("rot", np.eye(3)))}),
        with_elements(lambda e: {"locations": rng.uniform(-1, 1, np.asarray(e["locations"]).shape),
                                 "dn2": complex(rng.uniform(-1e-6, 1e-6), rng.uniform(-1e-8, 1e-8))}),
        {"elements": zero_el, "la": (1.0, 0.3, 0), "lb": (0, 1.0, 0.2), "lc": (0.5, 0.5, 1.0)},
        {"elements": zero_el, "rots": np.stack([_drz(a) @ _rx(b) for a, b in zip(model["_angles"], model["_tilts"])])},
        {"elements": zero_el, "centers": rng.uniform(-1, 1, (5, 3)), "weights": rng.uniform(-0.5, 0.5, 5),
         "dw_sigma": 1.0},
    ]
    return T


def _axpy(model, t, s):
    """model + s * tangent, field by field."""
    m = copy.deepcopy({k: v for k, v in model.items() if not k.startswith("_")})
    for e, te in zip(m["elements"], t.get("elements", [{}] * len(m["elements"]))):
        if "dims" in te:
            e["dims"] = [(np.asarray(v) + s * dv, np.asarray(w) + s * dw) for (v, w), (dv, dw) in zip(e["dims"], te["dims"])]
        for k in ("tau", "eta"):
            if k in te:
                e[k] = e.get(k, 0.0) + s * te[k]
        if "rot" in te:
            e["rot"] = np.asarray(e.get("rot", np.eye(3))) + s * te["rot"]
        if "locations" in te:
            e["locations"] = np.asarray(e["locations"], dtype=float) + s * te["locations"]
        if "dn2" in te:
            e["dn2"] = e["dn2"] + s * te["dn2"]
    for k in ("la", "lb", "lc"):
        if k in t:
            m[k] = tuple(np.asarray(m[k], dtype=float) + s * np.asarray(t[k]))
    for k in ("rots", "centers", "weights"):
        if k in t:
            m[k] = np.asarray(m[k]) + s * np.asarray(t[k])
    if "dw_sigma" in t:
        m["dw_sigma"] = m["dw_sigma"] + s * t["dw_sigma"]
    return m


@pytest.mark.parametrize("cell", sorted(CELLS))
@pytest.mark.parametrize("absorbing", [False, True])
def test_ensemble_tangent_matches_central_differences(gpu_ctx, oracle_mod, cell, absorbing):
    from hipgisaxs_amd.api import ensemble_intensity, ensemble_intensity_tangent

    rng = np.random.default_rng(21)
    g, fc = _grid(oracle_mod, gpu_ctx, absorbing)
    fcd = _dev(fc)
    model = _model(CELLS[cell], rng)
    T = _tangents(model, rng)
    I, dI = ensemble_intensity_tangent(gpu_ctx, {k: v for k, v in model.items() if not k.startswith("_")}, T, fc=fcd)
    I0 = ensemble_intensity(gpu_ctx, fc=fcd, **{k: v for k, v in model.items() if not k.startswith("_")})
    assert rel_l2(I.cpu().numpy(), I0.cpu().numpy()) <= 1e-12
    h = 1e-5
    for j, t in enumerate(T):
        ip = ensemble_intensity(gpu_ctx, fc=fcd, **_axpy(model, t, h)).cpu().numpy()
        im = ensemble_intensity(gpu_ctx, fc=fcd, **_axpy(model, t, -h)).cpu().numpy()
        cd = (ip - im) / (2 * h)
        e = rel_l2(dI[j].cpu().numpy(), cd)
        print(f"{cell} absorbing={absorbing} direction {j}: relL2 {e:.2e}")
        assert e <= 1e-6


@pytest.mark.parametrize("metric", ["chi2", "l2", "sqrt_unit_l2", "sqrt_c_l2", "cbrt_unit_l2", "cbrt_c_l2"])
def test_distance_tangent_matches_central_differences(gpu_ctx, metric):
    import torch

    from hipgisaxs_amd.api import image_distance, image_distance_tangent

    rng = np.random.default_rng(5)
    n = 5000
    I = rng.uniform(0.5, 50.0, n)
    O = rng.uniform(0.0, 60.0, n)
    O[::97] = 0.0
    mask = rng.uniform(0.2, 1.0, n)
    dI = rng.normal(size=(3, n))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    D, dD = image_distance_tangent(gpu_ctx, dev(I), dev(O), dev(mask), dev(dI), metric)
    d0 = image_distance(gpu_ctx, dev(I), dev(O), dev(mask), metric)
    assert D == pytest.approx(d0, rel=1e-14)
    h = 1e-4
    for j in range(3):
        dp = image_distance(gpu_ctx, dev(I + h * dI[j]), dev(O), dev(mask), metric)
        dm = image_distance(gpu_ctx, dev(I - h * dI[j]), dev(O), dev(mask), metric)
        cd = (dp - dm) / (2 * h)
        print(f"{metric} direction {j}: tangent {dD[j]:.10e} central {cd:.10e}")
        assert dD[j] == pytest.approx(cd, rel=1e-6, abs=1e-9 * abs(D))


# ----- the .hig fit -------------------------------------------------------------------------------

from pathlib import Path  # noqa: E402

GOLD = Path(__file__).resolve().parent / "golden" / "hig"
FIT = (GOLD / "fit_cylinders.hig").read_text()
TRUTH = [5.0, 50.0]
ALPHA_PARAM = """    fitparam = { key = "zalpha", variable = "instrumentation:scattering:alphai:min",
                 range = { min = 0.1, max = 0.3, step = 0.01 }, init = 0.175 },
    referencedata = {"""


@pytest.fixture(scope="module")
def fitdir(tmp_path_factory):
    import oracle.hig as OH
    from hipgisaxs_amd import hig as H

    d = tmp_path_factory.mktemp("fitgrad")
    params, _, _ = OH.read_fitting(FIT)
    m = OH.read_model(FIT)
    for p, v in zip(params, TRUTH):
        OH.update_param(m, p[1], v)
    H.HigSimulation.save(OH.run(None, model=m), str(d / "fit_cylinders_ref.out"))
    return d


def _oracle_grad(d, x, metric, reg=0.0, text=FIT, rel=1e-5):
    import oracle.hig as OH

    params, _, _ = OH.read_fitting(text)
    O = OH.read_ascii_image(str(d / "fit_cylinders_ref.out"))
    f = lambda y: OH.objective(text, params, y, metric, O, reg=reg)
    g = []
    for j in range(len(x)):
        h = rel * abs(x[j])
        xp, xm = list(x), list(x)
        xp[j] += h
        xm[j] -= h
        g.append((f(xp) - f(xm)) / (2 * h))
    return f(list(x)), np.array(g)


@pytest.mark.parametrize("metric,reg", [("sqrt_unit_l2", 0.0), ("sqrt_c_l2", 0.0), ("chi2", 0.0),
                                        ("cbrt_c_l2", 0.0), ("sqrt_unit_l2", 0.01)])
def test_hig_gradient_matches_oracle_central_differences(fitdir, metric, reg):
    """The tangent pass's gradient (radius: node values; lattice scaling: lattice vectors) against
    central differences of the oracle's objective (relative step 1e-5), to 1e-6 of the gradient."""
    from hipgisaxs_amd import hig as H

    sim = H.HigSimulation(text=FIT, base_dir=str(fitdir))
    x = [4.2, 48.5]
    f, g, done = sim.gradient(x, metric, reg)
    f0, g0 = _oracle_grad(fitdir, x, metric, reg)
    print(f"{metric} reg={reg}: f {f:.12e} (oracle {f0:.12e}), grad {g} (oracle {g0})")
    assert done.all()
    assert f == pytest.approx(f0, rel=1e-9)
    np.testing.assert_allclose(g, g0, rtol=0, atol=1e-6 * np.abs(g0).max())
    # the same value as the objective call
    assert sim.objective(x, metric, reg) == pytest.approx(f, rel=1e-12)


def test_hig_gradient_leaves_instrument_variables_to_the_caller(fitdir):
    """A fit variable on the incidence angle moves the q-grid: computed is False there (its component
    0), and the other components are unchanged."""
    from hipgisaxs_amd import hig as H

    text = FIT.replace("    referencedata = {", ALPHA_PARAM)
    sim = H.HigSimulation(text=text, base_dir=str(fitdir))
    keys = [p["key"] for p in sim.fit_params()]
    assert keys == ["radius", "scaling", "zalpha"]
    f, g, done = sim.gradient([4.2, 48.5, 0.175], "sqrt_unit_l2")
    assert list(done) == [True, True, False] and g[2] == 0.0
    sim2 = H.HigSimulation(text=FIT, base_dir=str(fitdir))
    f2, g2, _ = sim2.gradient([4.2, 48.5], "sqrt_unit_l2")
    assert f == pytest.approx(f2, rel=1e-12)
    np.testing.assert_allclose(g[:2], g2, rtol=1e-12)


def test_lmvm_on_tangent_passes_reaches_the_truth(fitdir):
    """LMVM with tangent-pass gradients ends at the truth with fewer objective evaluations than on
    central differences (HGX_HIG_FIT_FD_GRADIENT)."""
    from hipgisaxs_amd import hig as H

    fine = FIT.replace("step = 0.5 }", "step = 0.02 }").replace("step = 2.5 }", "step = 0.1 }")
    sim = H.HigSimulation(text=fine, base_dir=str(fitdir))
    out = sim.fit(algorithm=2)
    fd = H.HigSimulation(text=fine, base_dir=str(fitdir), flags=H.FIT_FD_GRADIENT).fit(algorithm=2)
    print(f"lmvm tangent: {out}; central differences: {fd}")
    np.testing.assert_allclose(out["x"], TRUTH, rtol=0, atol=0.02)
    assert out["evals"] < fd["evals"]


def test_pso_generation_batch_equals_particle_by_particle(fitdir):
    """One synchronisation per generation gives the particle-by-particle swarm exactly: the same
    best point, value, evaluation count and history."""
    from hipgisaxs_amd import hig as H

    a = H.HigSimulation(text=FIT, base_dir=str(fitdir))
    ra = a.fit(algorithm=0, seed=7)
    ha = a.fit_history()
    b = H.HigSimulation(text=FIT, base_dir=str(fitdir), flags=H.FIT_SEQUENTIAL)
    rb = b.fit(algorithm=0, seed=7)
    hb = b.fit_history()
    assert ra["evals"] == rb["evals"] == 24
    np.testing.assert_array_equal(ra["x"], rb["x"])
    assert ra["value"] == rb["value"]
    np.testing.assert_array_equal(ha, hb)
