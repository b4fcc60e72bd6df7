"""The .hig driver's host side (reading, scan configurations, grain sampling, mesh files, output
format) against the CPU oracle of the simulation loop and against the C library's rand(). No GPU."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

import oracle.hig as OH
from hipgisaxs_amd import hig as H
from hipgisaxs_amd.api import HgxError

GOLD = Path(__file__).resolve().parent / "golden" / "hig"
REF_EXAMPLES = Path("/root/reference/examples")

SAMPLER = """
hipGisaxsInput = {{
  shape = {{ key = "s", name = "sphere", param = {{ type = "radius", min = 4 }} }},
  unitcell = {{ key = "u", element = {{ shape:key = "s", locations = [[0 0 0]] }} }},
  structure = {{ key = "a", grain = {{ unitcell:key = "u" }},
                ensemble = {{ maxgrains = [7 5 3], spacing = [{sp}],
                             orientations = {{ stat = "{stat}",
                               rot1 = {{ axis = "z", angles = [0 360], anglemean = 100, anglesd = 40,
                                        anglelocation = 90, anglescale = 30 }},
                               rot2 = {{ axis = "x", angles = [0 10], anglemean = 2, anglesd = 3,
                                        anglelocation = 4, anglescale = 2 }},
                               rot3 = {{ axis = "{ax3}", angles = [5 20], anglemean = 9, anglesd = 4 }} }} }} }},
  structure = {{ key = "b", grain = {{ unitcell:key = "u", transvec = [1 2 3] }},
                ensemble = {{ maxgrains = [3 2 1], orientations = {{ stat = "{stat}",
                               rot1 = {{ axis = "y", angles = [10 30] }} }} }} }},
  instrumentation = {{ scattering = {{ inplanerot = {{ min = {phi} }} }} }},
  computation = {{ outputregion = {{ type = "qspace", minpoint = [-1 0], maxpoint = [1 1] }}, resolution = [8 6] }}
}}
"""


def test_normal_orientations_are_unweighted():
    """"normal" draws the angles and weights of "gaussian" but run_gisaxs weights only "gaussian" and
    "cauchy" grains."""
    g = H.HigSimulation(text=SAMPLER.format(stat="gaussian", ax3="z", sp="0 0 0", phi=0)).grains("a")
    n = H.HigSimulation(text=SAMPLER.format(stat="normal", ax3="z", sp="0 0 0", phi=0)).grains("a")
    np.testing.assert_array_equal(g[0], n[0])
    assert np.all(n[2] == 1.0) and not np.all(g[2] == 1.0)


def _grains_oracle(text, key, seed):
    m = OH.read_model(text)
    draws = OH.libc_rand(seed)
    phi = m["scattering"]["inplanerot"][0] * OH.PI_ / 180
    for k in sorted(m["structures"]):
        g = OH.grains(m["structures"][k], phi, draws)
        if k == key:
            return g
    raise KeyError(key)


@pytest.mark.parametrize("stat,ax3,sp", [("range", "y", "0 0 0"), ("random", "z", "0 0 0"),
                                         ("gaussian", "n", "40 30 20"), ("normal", "z", "0 0 0"),
                                         ("cauchy", "y", "5 5 0"), ("single", "x", "10 0 5")])
@pytest.mark.parametrize("seed", [1, 2, 987654, 3000000000])
def test_grains_match_oracle_and_libc_rand(stat, ax3, sp, seed):
    """Orientation sampling, grain centres and weights (spatial_distribution / orientation_distribution,
    src/sim/hipgisaxs_main.cpp:1702-2086) equal the oracle's, whose rand() is the C library's own."""
    text = SAMPLER.format(stat=stat, ax3=ax3, sp=sp, phi=12.5)
    sim = H.HigSimulation(text=text, seed=seed)
    for key in ("a", "b"):
        rot, cen, w = sim.grains(key)
        ref = _grains_oracle(text, key, seed)
        assert rot.shape[0] == len(ref)
        for i, (R, c, wt) in enumerate(ref):
            np.testing.assert_allclose(rot[i], R, rtol=0, atol=2e-16)
            np.testing.assert_allclose(cen[i], c, rtol=0, atol=1e-13)
            assert w[i] == pytest.approx(wt, rel=1e-15, abs=0)


def test_sampler_is_the_seed1_rand_sequence():
    """A reference process never seeds rand() on this path: with the other rotations at zero, grain k of
    a z-range [0, 360] structure turns by 360 * rand_k / RAND_MAX with rand_k the seed-1 sequence."""
    text = SAMPLER.format(stat="range", ax3="y", sp="0 0 0", phi=0)
    text = text.replace("angles = [0 10]", "angles = [0 0]").replace("angles = [5 20]", "angles = [0 0]")
    rot, _, _ = H.HigSimulation(text=text).grains("a")
    draws = OH.libc_rand(1)
    first = [next(draws) for _ in range(4)]
    assert first[:2] == [1804289383, 846930886]  # glibc's seed-1 sequence
    for k in range(4):
        a = math.degrees(math.atan2(rot[k][1, 0], rot[k][0, 0])) % 360
        assert a == pytest.approx(360.0 * first[k] / 2147483647.0, abs=1e-9)


def _fixture_texts():
    return sorted(GOLD.glob("*.hig"))


@pytest.mark.parametrize("path", _fixture_texts(), ids=lambda p: p.stem)
def test_fixture_model_matches_oracle(path):
    text = path.read_text()
    sim = H.HigSimulation(path=str(path))
    info = sim.info()
    m = OH.read_model(text)
    assert (info["ncol"], info["nrow"]) == tuple(m["compute"]["resolution"])
    sc = m["scattering"]
    keV = sc["photon"] / 1000 if sc["unit"] == "ev" else sc["photon"]
    assert info["k0"] == pytest.approx(2 * math.pi / (1.23984 / keV), rel=1e-15)
    for key in m["structures"]:
        rot, cen, w = sim.grains(key)
        ref = _grains_oracle(text, key, 1)
        np.testing.assert_allclose(rot, np.array([r[0] for r in ref]), atol=2e-16)
        np.testing.assert_allclose(cen, np.array([r[1] for r in ref]), atol=1e-13)


def test_scan_configurations():
    """run_all_gisaxs's loop: alpha_i outermost, counts (max - min) / step + 1 truncated; an alphai
    step also sets the in-plane step (the reference's switch falls through) unless inplanerot has one."""
    cfg = H.HigSimulation(path=str(GOLD / "two_structures.hig")).configurations()
    assert [(round(a, 12), round(p, 12)) for a, p, _ in cfg] == [(0.12, 0.0), (0.12, 15.0), (0.22, 0.0), (0.22, 15.0)]
    text = (GOLD / "spheres_bcc.hig").read_text().replace("alphai = { min = 0.175 }",
                                                          "alphai = { min = 0.1, max = 0.3, step = 0.05 }")
    cfg = H.HigSimulation(text=text).configurations()
    assert len(cfg) == int((0.3 - 0.1) / 0.05 + 1)
    # the fall-through: phi_step = 0.05 with phi_min = phi_max = 0 still yields one phi per alpha_i
    assert all(p == 0.0 for _, p, _ in cfg)
    single = H.HigSimulation(path=str(GOLD / "pyramid_dwba.hig")).configurations()
    assert single == [(0.16, 0.0, 0.0)]


@pytest.mark.skipif(not REF_EXAMPLES.exists(), reason="the reference's example inputs are not present")
@pytest.mark.parametrize("name", ["01-pyramid", "02-spheres-bcc", "03-cylinders-hcp", "07-gratings"])
def test_reference_examples_read(name):
    """The reference's own example inputs read into the same model as the oracle reads them."""
    path = REF_EXAMPLES / f"{name}.hig"
    text = path.read_text()
    sim = H.HigSimulation(path=str(path))
    info = sim.info()
    m = OH.read_model(text)
    assert (info["ncol"], info["nrow"]) == tuple(m["compute"]["resolution"])
    assert info["nconfig"] == 1
    for key in m["structures"]:
        rot, cen, w = sim.grains(key)
        ref = _grains_oracle(text, key, 1)
        assert rot.shape[0] == len(ref) == int(np.prod(m["structures"][key]["maxgrains"]))
        np.testing.assert_allclose(rot, np.array([r[0] for r in ref]), atol=2e-16)


@pytest.mark.parametrize("bad,needle", [
    ('hipGisaxsInput = { shape = { key = "s", name = "sphere", bogus = 1 } }', "bogus"),
    ('hipGisaxsInput = { shape = { key = "s", name = "sphere" ', "missing"),
    ('hipGisaxsInput = { shape = { key = "s", name = "blob" } }', "unknown shape name"),
    ('hipGisaxsInput = { unitcell = { key = "u", element = { shape:key = "s", locations = [[0 0]] } } }',
     "vector of 3"),
    ('hipGisaxsInput = { structure = { key = "s", iratio = 0 } }', "iratio"),
    ('hipGisaxsInput = { shape = { key = "s", name = "sphere", param = { type = "radius", stat = "cauchy" } } }',
     "not supported"),
    ('hipGisaxsInput = { computation = { resolution = [1 5] } }', "resolution"),
    ('hipGisaxsInput = { shape = { key = "s\n', "unterminated"),
    ('nothing = 1', "top-level"),
])
def test_read_errors_name_the_problem(bad, needle):
    with pytest.raises(HgxError, match=needle):
        H.HigSimulation(text=bad)


def test_unsupported_features_fail_at_run_not_silently():
    text = (GOLD / "spheres_bcc.hig").read_text().replace('maxgrains = [4 4 1]',
                                                          'maxgrains = [4 4 1], distribution = "random"')
    sim = H.HigSimulation(text=text)
    with pytest.raises(HgxError, match="random"):
        sim.grains("bcc")


def _one_grain_pair(corr):
    """The two-structure fixture with one grain per structure and the given structcorrelation."""
    t = (GOLD / "two_structures.hig").read_text()
    t = t.replace("maxgrains = [2 2 1]", "maxgrains = [1 1 1]").replace("maxgrains = [2 1 1]", "maxgrains = [1 1 1]")
    return t.replace("resolution = [40 24]", f'resolution = [40 24], structcorrelation = "{corr}"')


def test_oracle_structure_correlations_pin():
    """GE and GnE of the oracle through its own nGnE path: with iratios (3, 1) normalised to (3/4, 1/4),
    nGnE gives A = 3/4 I1 + 1/4 I2, and iratios (1, 0) give I1 with the same grain draws, so I2 follows;
    then GE = A^2 and GnE = 3/4 I1^2 + 1/4 I2^2, and more than one grain is refused."""
    A = OH.run(_one_grain_pair("nGnE"), smear_mode="off")
    I1 = OH.run(_one_grain_pair("nGnE").replace("iratio = 1,", "iratio = 0,"), smear_mode="off")
    I2 = (A - 0.75 * I1) / 0.25
    assert np.linalg.norm(I2) > 0.1 * np.linalg.norm(I1)
    ge = OH.run(_one_grain_pair("GE"), smear_mode="off")
    gne = OH.run(_one_grain_pair("GnE"), smear_mode="off")
    assert np.allclose(ge, A * A, rtol=1e-13, atol=0)
    assert np.allclose(gne, 0.75 * I1 * I1 + 0.25 * I2 * I2, rtol=1e-9, atol=1e-12 * np.abs(gne).max())
    with pytest.raises(ValueError):
        OH.run((GOLD / "two_structures.hig").read_text().replace(
            "resolution = [40 24]", 'resolution = [40 24], structcorrelation = "GE"'))


def _spheres(grain):
    """The spheres fixture with its grain's repetition and scaling replaced."""
    return (GOLD / "spheres_bcc.hig").read_text().replace("repetition = [100 100 10], scaling = 15", grain)


_SCALING3 = ('repetition = [20 20 4], scaling = { a = { stat = "gaussian", mean = 15, stddev = 0.3, nsamples = 3 }, '
             'b = { stat = "gaussian", mean = 15, stddev = 0.2, nsamples = 2 }, '
             'c = { stat = "cauchy", mean = 14, stddev = 0.25, nsamples = 3 } }')


def test_oracle_scaling_distribution_pins():
    """The oracle's grain scaling loop: every sample sits in its cell of [mean - 4 sd, mean + 4 sd]
    (8 sd / (n + 1) wide, n the largest sample count) with the pdf of the reference's constants as
    its weight, and the image is the sum over samples of weight^2 times the image at that fixed
    scaling; zero widths give the plain image."""
    st = OH.read_model(_spheres(_SCALING3))["structures"]["bcc"]
    sets = OH.grain_sets(st, [(None, None, 1.0)])
    assert len(sets) == 3
    means, sds = [15.0, 15.0, 14.0], [0.3, 0.2, 0.25]
    for j, (sc, rep, g) in enumerate(sets):
        assert rep == [20.0, 20.0, 4.0]
        for ax in range(3):
            lo, dx = means[ax] - 4 * sds[ax], 8 * sds[ax] / 4
            assert lo + j * dx <= sc[ax] < lo + (j + 1) * dx
        x = (sc[0] - 15.0) / 0.3
        assert g[0][2] == pytest.approx(math.exp(-0.5 * x * x) / (0.3 * 2.506628), rel=1e-14)
    img = OH.run(_spheres(_SCALING3), smear_mode="off")
    ref = sum(g[0][2] ** 2 * OH.run(_spheres(f"repetition = [20 20 4], scaling = [{sc[0]!r} {sc[1]!r} {sc[2]!r}]"),
                                     smear_mode="off") for sc, _, g in sets)
    assert np.allclose(img, ref, rtol=1e-12, atol=0)
    zero = _SCALING3.replace("stddev = 0.25", "stddev = 0").replace("stddev = 0.3", "stddev = 0").replace(
        "stddev = 0.2", "stddev = 0").replace("mean = 14", "mean = 15")
    assert np.allclose(OH.run(_spheres(zero), smear_mode="off"),
                       OH.run(_spheres("repetition = [20 20 4], scaling = 15"), smear_mode="off"), rtol=1e-14, atol=0)


def test_oracle_repetition_range_pins():
    """A repetition range of one value per axis gives the fixed repetitions; without a scaling
    distribution the reference's loop runs over three repetition entries (its sample list has three
    values) and adds each grain three times."""
    rd = ("scaling = 15, repetitiondist = { xrepetition = { stat = \"range\", min = 7, max = 7 }, "
          "yrepetition = { stat = \"range\", min = 6, max = 6 }, zrepetition = { stat = \"range\", min = 3, max = 3 } }")
    img = OH.run(_spheres(rd), smear_mode="off")
    ref = OH.run(_spheres("repetition = [7 6 3], scaling = 15"), smear_mode="off")
    assert np.allclose(img, 3 * ref, rtol=1e-13, atol=0)


def test_read_mesh_files():
    kind, tri = H.read_mesh(str(GOLD / "tetra.obj"))
    assert kind == "exact9" and tri.shape == (4, 9)
    np.testing.assert_array_equal(tri[3], [10, 0, 0, 0, 10, 0, 0, 0, 10])  # "f -3 -2 -1": the last three vertices
    # the quadrilateral cube splits into 12 triangles enclosing its volume
    kind, tri = H.read_mesh(str(GOLD / "cube_quads.obj"))
    assert tri.shape == (12, 9)
    v = tri.reshape(-1, 3, 3)
    vol = np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6
    assert vol == pytest.approx(64.0, rel=1e-15)
    kind, rec = H.read_mesh(str(GOLD / "two_plates.dat"))
    assert kind == "approx8"
    np.testing.assert_array_equal(rec, [[100, 0, 0, -1, 1, 1, 0, 0], [100, 0, 0, 1, 1, 1, 4, 0]])
    with pytest.raises(HgxError):
        H.read_mesh(str(GOLD / "missing.obj"))


def test_save_format(tmp_path):
    """save_gisaxs: nrow lines of ncol tab-terminated values, row-major."""
    img = np.arange(12, dtype=float).reshape(3, 4) / 7.0
    p = tmp_path / "out.txt"
    H.HigSimulation.save(img, str(p))
    lines = p.read_text().splitlines()
    assert len(lines) == 3 and all(line.endswith("\t") for line in lines)
    back = np.array([[float(x) for x in line.split("\t") if x] for line in lines])
    np.testing.assert_array_equal(back, img)


def test_oracle_smear_pins():
    """The oracle's smearing: weights normalised (a constant image is unchanged); in the interior the
    symmetric form is the plain discrete Gaussian convolution and the reference's window drops the
    k = i + h term only (include/numerics/convolutions.hpp:141-147)."""
    rng = np.random.default_rng(5)
    img = rng.random((20, 30))
    const = np.full((9, 11), 2.5)
    for compat in (False, True):
        np.testing.assert_allclose(OH.smear(const, 1.0, compat), const, rtol=1e-15)
    sig, h = 1.3, math.ceil(6 * 1.3)
    g = np.exp(-np.arange(-h, h + 1) ** 2 / (2 * sig * sig))
    out = OH.smear(img, sig)
    i, j = 10, 15  # interior in both directions (h = 8 < 10)
    rows = np.array([np.dot(img[r, j - h:j + h + 1], g) / g.sum() for r in range(i - h, i + h + 1)])
    assert out[i, j] == pytest.approx(np.dot(rows, g) / g.sum(), rel=1e-13)
    line = img[3]
    one_fixed = OH.smear(line[None, :], sig)  # one row: the vertical pass sees a single sample
    gl = g[:-1]
    assert one_fixed[0, 15] == pytest.approx(np.dot(line[15 - h:15 + h + 1], g) / g.sum(), rel=1e-13)
    one_compat = OH.smear(line[None, :], sig, compat=True)
    assert one_compat[0, 15] == pytest.approx(np.dot(line[15 - h:15 + h], gl) / gl.sum(), rel=1e-13)


def _read_tiff16(path):
    """Minimal reader of the driver's baseline TIFF: (width, height, uint16 pixels, tags)."""
    import struct
    b = Path(path).read_bytes()
    assert b[:4] == b"II*\x00"
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    tags = {}
    for k in range(n):
        tag, typ, cnt, val = struct.unpack_from("<HHII", b, ifd + 2 + 12 * k)
        tags[tag] = val & 0xFFFF if typ == 3 else val
    w, h = tags[256], tags[257]
    px = np.frombuffer(b, dtype="<u2", count=w * h, offset=tags[273]).reshape(h, w)
    return w, h, px, tags


def test_tiff_image_follows_the_reference_transform(tmp_path):
    """Image::construct_image: shift to non-negative, log10 with zeros kept, min-max normalisation; here
    16-bit grayscale. Closed form on (0, 1, 10, 100, 1000, 10): logs (0, 0, 1, 2, 3, 1) -> (0, 0, 1/3,
    2/3, 1, 1/3)."""
    img = np.array([[0.0, 1.0, 10.0], [100.0, 1000.0, 10.0]])
    H.HigSimulation.save_tiff(img, tmp_path / "a.tif")
    w, h, px, tags = _read_tiff16(tmp_path / "a.tif")
    assert (w, h, tags[258], tags[259], tags[262], tags[277]) == (3, 2, 16, 1, 1, 1)
    third = round(65535 / 3)
    np.testing.assert_array_equal(px, [[0, 0, third], [round(2 * 65535 / 3), 65535, third]])
    # a negative minimum is shifted away first: (-1, 0, 9) -> (0, 1, 10) -> logs (0, 0, 1)
    H.HigSimulation.save_tiff(np.array([[-1.0, 0.0, 9.0]]), tmp_path / "b.tif")
    np.testing.assert_array_equal(_read_tiff16(tmp_path / "b.tif")[2], [[0, 0, 65535]])
    # a constant image maps to 1 (0 when its log is negative); non-finite pixels are refused
    H.HigSimulation.save_tiff(np.full((2, 2), 5.0), tmp_path / "c.tif")
    assert np.all(_read_tiff16(tmp_path / "c.tif")[2] == 65535)
    with pytest.raises(Exception, match="non-finite"):
        H.HigSimulation.save_tiff(np.array([[1.0, np.nan]]), tmp_path / "d.tif")
