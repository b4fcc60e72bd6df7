"""Fitting a .hig model on the device against the oracle: the distance functors of hgx_image_distance,
the objective of HipGISAXSObjectiveFunction (update_params, compute_gisaxs, distance, regularization),
the particle-swarm and brute-force drivers step by step, LMVM convergence, and the override of the
output region by the reference data's region."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle.hig as OH
from oracle import metrics as OM
from hipgisaxs_amd import _native as N
from hipgisaxs_amd import api
from hipgisaxs_amd import hig as H

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "hig"
FIT = (GOLD / "fit_cylinders.hig").read_text()
TRUTH = [5.0, 50.0]


def _say(*a):
    print("[fit]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fitdir(tmp_path_factory):
    """The fit input beside reference data that the oracle computes at the true parameters."""
    d = tmp_path_factory.mktemp("fit")
    params, _, _ = OH.read_fitting(FIT)
    m = OH.read_model(FIT)
    for p, v in zip(params, TRUTH):
        OH.update_param(m, p[1], v)
    ref = OH.run(None, model=m)
    H.HigSimulation.save(ref, str(d / "fit_cylinders_ref.out"))
    (d / "fit.hig").write_text(FIT)
    return d


def _sim(d, text=FIT):
    return H.HigSimulation(text=text, base_dir=str(d))


def _oracle_f(d, x, metric, reg=0.0, text=FIT):
    params, _, _ = OH.read_fitting(text)
    O = OH.read_ascii_image(str(d / "fit_cylinders_ref.out"))
    return OH.objective(text, params, x, metric, O, reg=reg)


def test_image_distance_matches_oracle_functors():
    _say("image_distance")
    ctx = api.Context()
    rng = np.random.default_rng(5)
    for n in (1, 257, 100003):
        I = rng.gamma(2.0, 3.0, n)
        O = rng.poisson(I).astype(float) + 0.5
        m = (rng.uniform(size=n) > 0.2).astype(float) * rng.choice([1.0, 2.0], n)
        dI, dO, dm = (torch.tensor(a, dtype=torch.float64, device="cuda") for a in (I, O, m))
        for name, fn in OM.METRICS.items():
            got = api.image_distance(ctx, dI, dO, dm, metric=name)
            want = fn(O, I, m)
            _say(n, name, got, want)
            # one pixel: the unit-vector forms are zero up to rounding
            assert got == pytest.approx(want, rel=1e-11, abs=1e-20)
        # no mask: all weights 1
        assert api.image_distance(ctx, dI, dO, None, "sqrt_c_l2") == pytest.approx(OM.sqrt_c_l2(O, I, np.ones(n)),
                                                                                   rel=1e-11)
    # a NaN model pixel: L2 and chi^2 keep the term (NaN, so the objective takes its 1e8 penalty), the
    # square-root forms drop it, as the oracle's functors
    I = np.array([1.5, np.nan, 2.0, 4.0])
    O = np.array([1.0, 2.0, 3.0, 5.0])
    dI, dO = (torch.tensor(a, dtype=torch.float64, device="cuda") for a in (I, O))
    assert np.isnan(api.image_distance(ctx, dI, dO, None, "l2"))
    assert np.isnan(api.image_distance(ctx, dI, dO, None, "chi2"))
    assert api.image_distance(ctx, dI, dO, None, "sqrt_c_l2") == pytest.approx(OM.sqrt_c_l2(O, I, np.ones(4)),
                                                                               rel=1e-11)


@pytest.mark.parametrize("metric", ["sqrt_unit_l2", "sqrt_c_l2", "cbrt_unit_l2", "cbrt_c_l2", "chi2", "l2"])
def test_objective_matches_oracle(fitdir, metric):
    sim = _sim(fitdir)
    for x, reg in (([4.2, 48.5], 0.0), ([6.5, 58.0], 0.3), ([5.0, 50.0], 0.0)):
        got = sim.objective(x, metric, reg)
        want = _oracle_f(fitdir, x, metric, reg)
        _say(metric, x, reg, got, want)
        if x == TRUTH and reg == 0.0:
            assert abs(got) <= 1e-12 * max(1.0, abs(_oracle_f(fitdir, [4.2, 48.5], metric)))
        else:
            assert got == pytest.approx(want, rel=1e-9)


def test_pso_follows_the_oracle_swarm(fitdir):
    """The swarm's positions generation by generation equal the oracle's with the same mt19937_64 seed;
    the objective values agree to rounding, so every best update takes the same branch."""
    sim = _sim(fitdir)
    out = sim.fit(algorithm=0, seed=7)
    rows = sim.fit_history()
    params, _, algos = OH.read_fitting(FIT)
    a = algos[0]["params"]
    best, fb, ref_rows = OH.pso(lambda x: _oracle_f(fitdir, x, "sqrt_unit_l2"), [p[2] for p in params],
                                [p[3] for p in params], a["pso_omega"], a["pso_phi1"], a["pso_phi2"],
                                int(a["pso_num_particles"]), int(a["pso_num_generations"]), seed=7)
    ref_rows = np.array(ref_rows)
    _say("pso", out, "oracle best", best, fb)
    assert out["evals"] == len(ref_rows) == 24
    np.testing.assert_allclose(rows[:, :2], ref_rows[:, :2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rows[:, 2], ref_rows[:, 2], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(out["x"], best, rtol=0, atol=1e-12)
    assert out["value"] == pytest.approx(fb, rel=1e-9)
    # another seed moves the swarm differently
    assert not np.allclose(sim.fit(algorithm=0, seed=8)["x"], out["x"])


def test_bruteforce_visits_the_grid_and_finds_the_truth(fitdir):
    sim = _sim(fitdir)
    out = sim.fit(algorithm=1)
    rows = sim.fit_history()
    params, _, _ = OH.read_fitting(FIT)
    pts = np.array(OH.bruteforce_points([p[2] for p in params], [p[3] for p in params], [p[4] for p in params]))
    _say("bruteforce", out)
    assert rows.shape[0] == len(pts) == 99
    np.testing.assert_array_equal(rows[:, :2], pts)
    np.testing.assert_allclose(out["x"], TRUTH, rtol=0, atol=1e-12)
    assert out["value"] <= 1e-12
    # spot-check three grid values against the oracle objective
    for k in (0, 17, 98):
        assert rows[k, 2] == pytest.approx(_oracle_f(fitdir, list(pts[k]), "sqrt_c_l2"), rel=1e-9)


def test_lmvm_converges_to_the_truth(fitdir):
    """From initial values inside the truth's basin (the lattice scaling has side minima a few units
    away), LMVM on central differences over the fitparam steps: with the brute-force grid's coarse steps
    it descends by more than 20x; with fine steps it ends at the truth."""
    sim = _sim(fitdir)
    f0 = sim.objective([4.2, 48.5], "sqrt_unit_l2")
    out = sim.fit(algorithm=2)
    _say("lmvm coarse steps", out, "f0", f0)
    assert out["value"] < 0.05 * f0
    assert 3 <= out["x"][0] <= 8 and 40 <= out["x"][1] <= 60
    fine = FIT.replace("step = 0.5 }", "step = 0.02 }").replace("step = 2.5 }", "step = 0.1 }")
    sim = _sim(fitdir, fine)
    out = sim.fit(algorithm=2)
    _say("lmvm fine steps", out)
    assert out["value"] < 1e-6 * f0
    np.testing.assert_allclose(out["x"], TRUTH, rtol=0, atol=0.02)


def test_lmvm_regularization_continuation_and_fp32c_objective(fitdir):
    """A regularized LMVM runs the reference's continuation (factor / 5 per solve, warm starts) and ends
    at the truth; the FP32C simulation gives the objective within the FP32C image bound."""
    fine = FIT.replace("step = 0.5 }", "step = 0.02 }").replace("step = 2.5 }", "step = 0.1 }")
    fine = fine.replace("tolerance = 1e-10 }", "tolerance = 1e-10, regularization = 0.01 }")
    sim = _sim(fitdir, fine)
    assert sim.fit_algorithms()[2]["regularization"] == 0.01
    out = sim.fit(algorithm=2)
    _say("lmvm regularized", out)
    np.testing.assert_allclose(out["x"], TRUTH, rtol=0, atol=0.02)
    # the last solve's factor is 0.01 / 5^9: the value is the distance plus that small penalty
    assert out["value"] == pytest.approx(sim.objective(out["x"], "sqrt_unit_l2", 0.01 / 5 ** 9), rel=1e-9)
    x = [4.2, 48.5]
    f64 = _sim(fitdir).objective(x, "sqrt_c_l2")
    s32 = H.HigSimulation(text=FIT, base_dir=str(fitdir), precision="fp32c")
    f32 = s32.objective(x, "sqrt_c_l2")
    _say("fp32c objective", f32, f64)
    assert f32 == pytest.approx(f64, rel=1e-5)


def test_plain_objective_algorithm(fitdir):
    sim = _sim(fitdir)
    out = sim.fit(algorithm=3)
    assert out["evals"] == 1
    np.testing.assert_array_equal(out["x"], [4.2, 48.5])
    assert out["value"] == pytest.approx(_oracle_f(fitdir, [4.2, 48.5], "cbrt_unit_l2"), rel=1e-9)


def test_reference_region_overrides_the_output_region(tmp_path):
    """Reference data of another size: the simulation takes the reference's region and size
    (override_qregion) and the objective equals the oracle's on that region."""
    text = FIT.replace("regmin = [-1 0], regmax = [1 1]", "regmin = [-0.6 0.1], regmax = [0.8 0.9]")
    params, _, _ = OH.read_fitting(text)
    m = OH.read_model(text)
    m["compute"]["minpoint"], m["compute"]["maxpoint"] = [-0.6, 0.1], [0.8, 0.9]
    m["compute"]["resolution"] = [30, 20]
    for p, v in zip(params, TRUTH):
        OH.update_param(m, p[1], v)
    ref = OH.run(None, model=m)
    H.HigSimulation.save(ref, str(tmp_path / "fit_cylinders_ref.out"))
    sim = H.HigSimulation(text=text, base_dir=str(tmp_path))
    sim.fit_init()
    assert (sim.info()["ncol"], sim.info()["nrow"]) == (30, 20)
    x = [6.0, 44.0]
    got = sim.objective(x, "sqrt_c_l2")
    m2 = OH.read_model(text)
    m2["compute"] = m["compute"]
    for p, v in zip(params, x):
        OH.update_param(m2, p[1], v)
    want = OM.sqrt_c_l2(ref.ravel(), OH.run(None, model=m2).ravel(), np.ones(ref.size))
    _say("override", got, want)
    assert got == pytest.approx(want, rel=1e-9)
    assert sim.objective(TRUTH, "sqrt_c_l2") <= 1e-12


def _write_edf(path, img):
    rows, cols = img.shape
    items = ["HeaderID = EH:000001:000000:000000", "ByteOrder = LowByteFirst", "DataType = FloatValue",
             f"Dim_1 = {cols}", f"Dim_2 = {rows}", f"EDF_BinarySize = {4 * rows * cols}"]
    head = "{\n" + "".join(f"{it} ;\n" for it in items)
    head += " " * (1024 - len(head) - 2) + "}\n"
    Path(path).write_bytes(head.encode() + np.asarray(img, "<f4").tobytes())


def test_edf_reference_data(fitdir, tmp_path):
    """Reference data in an ESRF data file (EDFReader: float32 pixels after the header blocks) give the
    objective of the same float32 values written as ASCII, and the oracle's with its EDF reader."""
    O32 = OH.read_ascii_image(str(fitdir / "fit_cylinders_ref.out")).astype(np.float32)
    _write_edf(tmp_path / "ref.edf", O32)
    np.savetxt(tmp_path / "ref.out", O32.astype(np.float64), fmt="%.17g")
    t_edf = FIT.replace("fit_cylinders_ref.out", "ref.edf")
    t_out = FIT.replace("fit_cylinders_ref.out", "ref.out")
    assert t_edf != FIT
    s_edf = H.HigSimulation(text=t_edf, base_dir=str(tmp_path))
    s_out = H.HigSimulation(text=t_out, base_dir=str(tmp_path))
    params, _, _ = OH.read_fitting(t_edf)
    O = OH.read_edf_image(str(tmp_path / "ref.edf"))
    x = [5.5, 49.0]
    for metric in ("sqrt_unit_l2", "chi2"):
        a, b = s_edf.objective(x, metric), s_out.objective(x, metric)
        want = OH.objective(t_edf, params, x, metric, O)
        _say("edf", metric, a, b, want)
        assert a == pytest.approx(b, rel=1e-12)
        assert a == pytest.approx(want, rel=1e-9)


def test_mask_file_weights_the_distance(fitdir, tmp_path):
    """A referencedata mask (whitespace-separated weights, read_mask_data) weights every functor as the
    oracle's mask does; a zero mask over the lower half changes the distance."""
    import shutil
    shutil.copy(fitdir / "fit_cylinders_ref.out", tmp_path / "fit_cylinders_ref.out")
    mask = np.ones((24, 40))
    mask[12:, :] = 0
    mask[:3, :5] = 2
    np.savetxt(tmp_path / "mask.txt", mask, fmt="%d")
    text = FIT.replace('mask = ""', 'mask = "mask.txt"')
    sim = H.HigSimulation(text=text, base_dir=str(tmp_path))
    params, _, _ = OH.read_fitting(text)
    O = OH.read_ascii_image(str(tmp_path / "fit_cylinders_ref.out"))
    x = [6.0, 47.5]
    for metric in ("sqrt_unit_l2", "cbrt_c_l2", "chi2"):
        got = sim.objective(x, metric)
        want = OH.objective(text, params, x, metric, O, mask=mask)
        _say("mask", metric, got, want)
        assert got == pytest.approx(want, rel=1e-9)
        assert abs(got - _oracle_f(fitdir, x, metric)) > 1e-6 * abs(got)
