"""Host side of fitting a .hig model (no device work): the fitting block as the driver reads it
against the oracle's reading, the parameter-path updates of HiGInput::update_params (their effect on
the sampled grains, and the paths both refuse), the particle-swarm generator and the brute-force grid
of the oracle."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle.hig as OH
from hipgisaxs_amd import _native as N
from hipgisaxs_amd import hig as H
from hipgisaxs_amd.api import HgxError

GOLD = Path(__file__).resolve().parent / "golden" / "hig"
FIT = (GOLD / "fit_cylinders.hig").read_text()


def test_fitting_block_matches_oracle_reading():
    sim = H.HigSimulation(text=FIT, base_dir=str(GOLD))
    params, ref, algos = OH.read_fitting(FIT)
    got = sim.fit_params()
    assert [p["key"] for p in got] == ["radius", "scaling"]  # the fit vector is ordered by key
    for g, p in zip(got, params):
        assert (g["key"], g["min"], g["max"], g["step"], g["init"]) == (p[0], p[2], p[3], p[4], p[5])
    ga = sim.fit_algorithms()
    assert [a["name"] for a in ga] == [a["name"] for a in algos] == ["pso", "bruteforce", "lmvm", "none_pounders"]
    assert [a["metric"] for a in ga] == [N.METRICS[a["metric"]] for a in algos]
    assert ga[2]["tolerance"] == 1e-10 and ref["regmax"] == [1.0, 1.0]


def test_fitparam_without_variable_or_duplicate_key_is_refused():
    bad = FIT.replace('variable = "shape[\'cyl\']:param[\'radius\']:min",', "")
    with pytest.raises(HgxError, match="incomplete fit parameter"):
        H.HigSimulation(text=bad)
    dup = FIT.replace('key = "scaling", variable', 'key = "radius", variable')
    with pytest.raises(HgxError, match="duplicate key"):
        H.HigSimulation(text=dup)


SAMPLER = """
hipGisaxsInput = {
  shape = { key = "s", name = "sphere", param = { type = "radius", min = 4 } },
  unitcell = { key = "c", element = { shape:key = "s", locations = [[0 0 0]] } },
  layer = { key = "film", order = 1, thickness = 20, refindex = { delta = 1e-6, beta = 1e-8 } },
  structure = { key = "a",
    grain = { unitcell:key = "c", lattice = { type = "bcc" }, repetition = [3 3 1], transvec = [1 2 3] },
    ensemble = { maxgrains = [2 2 1], spacing = [40 30 0],
                 orientations = { stat = "range", rot1 = { axis = "z", angles = [0 90] } } } },
  instrumentation = { scattering = { alphai = { min = 0.2 }, inplanerot = { min = 7.5 },
                      photon = { value = 10000, unit = "ev" } } },
  computation = { outputregion = { type = "qspace", minpoint = [-1 0], maxpoint = [1 1] }, resolution = [8 6] }
}
"""


@pytest.mark.parametrize("updates", [
    [("structure['a']:ensemble:maxgrains:x", 3), ("structure['a']:ensemble:spacing:y", 55.0)],
    [("structure['a']:grain:transvec:z", -4.5), ("structure['a']:ensemble:maxgrains:2", 2)],
    [("instrumentation:scattering:inplanerot:min", 33.0)],
])
def test_updates_reach_the_sampled_grains(updates):
    """Grain counts, spacings, translations and the in-plane angle set by update_param give the grains the
    oracle samples from the same edited model."""
    sim = H.HigSimulation(text=SAMPLER)
    m = OH.read_model(SAMPLER)
    before = sim.grains("a")[1].copy()
    for var, v in updates:
        sim.update_param(var, v)
        OH.update_param(m, var, v)
    rot, cen, w = sim.grains("a")
    phi = m["scattering"]["inplanerot"][0] * OH.PI_ / 180
    ref = OH.grains(m["structures"]["a"], phi, OH.libc_rand(1))
    assert rot.shape[0] == len(ref)
    for i, (R, c, wt) in enumerate(ref):
        np.testing.assert_allclose(rot[i], R, rtol=0, atol=2e-16)
        np.testing.assert_allclose(cen[i], c, rtol=0, atol=1e-12)
    assert cen.shape != before.shape or not np.allclose(cen, before)


GOOD_PATHS = [
    "shape['s']:param['radius']:min", "shape['s']:param['radius']:p1", "shape['s']:param['radius']:nvalues",
    "shape['s']:zrot", "shape['s']:originvec:y", "layer['film']:thickness", "layer['film']:refindex:delta",
    "structure['a']:iratio", "structure['a']:grain:lattice:a:x", "structure['a']:grain:lattice:c:2",
    "structure['a']:grain:scaling", "structure['a']:grain:scaling:b:mean", "structure['a']:grain:repetition:y",
    "instrumentation:scattering:alphai:min", "instrumentation:scattering:photon:value",
]
BAD_PATHS = [
    "shape['nope']:param['radius']:min", "shape['s']:param['height']:min", "shape['s']:param['radius']:mean",
    "shape[s]:zrot", "shape['s'", "layer['film']:refindex:gamma", "layer['x']:thickness",
    "structure['a']:grain:lattice:a:w", "structure['a']:grain:bogus", "structure['a']:ensemble:spacing:q",
    "instrumentation:scattering:alphai:mean", "nothing['a']:x",
]


def test_update_paths_accepted_and_refused_alike():
    for var in GOOD_PATHS:
        sim = H.HigSimulation(text=SAMPLER)
        sim.update_param(var, 2.0)
        OH.update_param(OH.read_model(SAMPLER), var, 2.0)
    for var in BAD_PATHS:
        sim = H.HigSimulation(text=SAMPLER)
        with pytest.raises(HgxError):
            sim.update_param(var, 2.0)
        if "[" in var and "']" not in var.split(":")[0]:
            continue  # malformed brackets: the oracle's splitter reads them as plain names
        with pytest.raises((KeyError, StopIteration)):
            OH.update_param(OH.read_model(SAMPLER), var, 2.0)


def test_lattice_component_update_starts_from_the_type_vectors():
    """Setting one component of a type-built lattice keeps the type's other vectors (bcc c = (.5 .5 .5))."""
    m = OH.read_model(SAMPLER)
    OH.update_param(m, "structure['a']:grain:lattice:a:y", 0.25)
    assert m["structures"]["a"]["abc"][0] == [1.0, 0.25, 0.0]
    assert m["structures"]["a"]["abc"][2] == [0.5, 0.5, 0.5]


def test_fit_init_reports_missing_reference_data(tmp_path):
    sim = H.HigSimulation(text=FIT, base_dir=str(tmp_path))
    with pytest.raises(HgxError, match="unable to open reference data"):
        sim.fit_init()
    edf = FIT.replace("fit_cylinders_ref.out", "ref.edf")
    with pytest.raises(HgxError, match="EDF"):
        H.HigSimulation(text=edf, base_dir=str(tmp_path)).fit_init()


def _write_edf(path, img, header_blocks=1, first_key="HeaderID = EH:000001:000000:000000"):
    """An ESRF data file: a '{ key = value ; ... }' header padded with spaces to header_blocks x 512
    bytes, then the image as little-endian float32, row by row."""
    rows, cols = img.shape
    items = [first_key, "Image = 1", "ByteOrder = LowByteFirst", "DataType = FloatValue",
             f"Dim_1 = {cols}", f"Dim_2 = {rows}", f"Size = {4 * rows * cols}", f"EDF_BinarySize = {4 * rows * cols}"]
    head = "{\n" + "".join(f"{it} ;\n" for it in items)
    head += " " * (512 * header_blocks - len(head) - 2) + "}\n"
    assert len(head) == 512 * header_blocks
    Path(path).write_bytes(head.encode() + np.asarray(img, "<f4").tobytes())


def test_oracle_edf_reader_pins(tmp_path):
    """The oracle's EDF reader returns the float32 image a file holds, for a header of one or of two
    512-byte blocks, with an item split across the block boundary."""
    rng = np.random.default_rng(3)
    img = rng.gamma(2.0, 50.0, (7, 11)).astype(np.float32)
    _write_edf(tmp_path / "a.edf", img)
    got = OH.read_edf_image(str(tmp_path / "a.edf"))
    assert got.shape == (7, 11) and np.array_equal(got, img.astype(np.float64))
    # a long first item pushes Dim_* across the first block boundary
    _write_edf(tmp_path / "b.edf", img, header_blocks=2, first_key="HeaderID = " + "X" * 470)
    assert np.array_equal(OH.read_edf_image(str(tmp_path / "b.edf")), img.astype(np.float64))


def test_mersenne_twister_64_is_the_standard_generator():
    """[rand.predef]: the 10000th output of a default-constructed std::mt19937_64 is
    9981545732273789042; the uniform draw is (g - min) / (max - min) in double."""
    g = OH.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042
    u = [OH.MT19937_64(7).uniform() for _ in range(3)]
    assert all(0.0 <= x <= 1.0 for x in u) and u[0] == u[1] == u[2]


def test_pso_oracle_minimises_a_quadratic():
    f = lambda x: (x[0] - 1.5) ** 2 + (x[1] + 0.5) ** 2
    best, fb, rows = OH.pso(f, [-3, -3], [3, 3], 0.6, 1.4, 1.4, 12, 30, seed=3)
    assert len(rows) == 360 and fb < 1e-4
    assert best == pytest.approx([1.5, -0.5], abs=1e-2)
    # positions never leave the limits
    assert np.all(np.abs(np.array(rows)[:, :2]) <= 3.0)


def test_bruteforce_grid_accumulates_its_step():
    """p += step as the reference loops: 0.1 added three times exceeds 0.3, so [0, 0.3] gives 3 points."""
    assert OH.bruteforce_points([0.0], [0.3], [0.1]) == [[0.0], [0.1], [0.2]]
    pts = OH.bruteforce_points([3, 40], [8, 60], [0.5, 2.5])
    assert len(pts) == 11 * 9 and pts[:2] == [[3, 40], [3, 42.5]]
    assert OH.bruteforce_points([1], [3], [-1]) == [[1], [2], [3]]  # no step: 1


def test_mask_of_the_wrong_size_is_refused(tmp_path):
    (tmp_path / "fit_cylinders_ref.out").write_text("\n".join(" ".join(["1.0"] * 40) for _ in range(24)) + "\n")
    (tmp_path / "mask.txt").write_text(" ".join(["1"] * 100))
    text = FIT.replace('mask = ""', 'mask = "mask.txt"')
    with pytest.raises(HgxError, match="mask and reference data dimension sizes do not match"):
        H.HigSimulation(text=text, base_dir=str(tmp_path)).fit_init()


@pytest.mark.parametrize("bad", [-2.0, 0.0, 1.5, 2e6])
def test_maxgrains_update_out_of_range_is_refused(bad):
    """maxgrains takes the parser's rule (an integer from 1 to 1e6 per axis) in update_param too: a
    negative pair of axes used to pass the product check and abort the process in make_grains."""
    sim = H.HigSimulation(text=SAMPLER)
    with pytest.raises(HgxError, match="maxgrains"):
        sim.update_param("structure['a']:ensemble:maxgrains:x", bad)
    with pytest.raises(HgxError, match="maxgrains"):
        sim.update_param("structure['a']:ensemble:maxgrains:y", -1.0)
    # the model is unchanged and still samples its grains
    rot, cen, w = sim.grains("a")
    assert rot.shape[0] == 4
