"""The .hig driver on the device (libhgx_hig.so over libhgx.so) against the CPU oracle of the
simulation loop (oracle/hig.py over the FP64 oracle kernels): whole images of the committed .hig
fixtures, every path of the driver (fused grain-ensemble kernel, per-grain calls, mesh shapes, SAXS,
scans of incidence and in-plane angles, the reference-compatible switches)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import oracle.hig as OH
from hipgisaxs_amd import hig as H

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "hig"
TOL64 = 1e-10   # FP64 driver vs FP64 oracle, relative L2 over the image
TOL32 = 1e-5    # FP32C paths


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _say(*a):
    print("[hig]", *a, flush=True)
    sys.stdout.flush()


def _text(name):
    return (GOLD / f"{name}.hig").read_text()


@pytest.mark.parametrize("flags", [0, H.NO_FUSE], ids=["default", "per_grain"])
@pytest.mark.parametrize("name", ["pyramid_dwba", "spheres_bcc", "cylinders_hcp", "box_grating", "saxs_cubes",
                                  "mesh_tetra"])
def test_fixture_image_matches_oracle(name, flags):
    _say(name, "flags", flags)
    text = _text(name)
    sim = H.HigSimulation(text=text, base_dir=str(GOLD), flags=flags)
    img = sim.compute()
    counts = sim.last_counts()
    ref = OH.run(text, base_dir=str(GOLD))
    assert img.shape == ref.shape
    assert np.all(np.isfinite(img)) and np.linalg.norm(ref) > 0
    r = _rel(img, ref)
    _say(name, "relL2", r, counts)
    assert r <= TOL64
    fusable = name != "mesh_tetra"  # analytic cells run fused (lattice or ensemble kernel)
    assert (counts["fused_grains"] > 0) == (fusable and flags == 0)


def test_scan_of_two_structures_matches_oracle():
    """Two structures (iratio 3:1), a two-element cell with three locations, shape rotations, a gaussian
    size distribution, gaussian-weighted orientations with grain spacing, a film layer, and a scan of
    2 x 2 (alpha_i, phi): each run draws its orientations after the previous ones, as the reference's
    run_all_gisaxs does."""
    text = _text("two_structures")
    sim = H.HigSimulation(text=text)
    imgs = sim.run_all()
    assert len(imgs) == 4
    for c, img in enumerate(imgs):
        ref = OH.run(text, config=c)
        r = _rel(img, ref)
        _say("two_structures config", c, "relL2", r)
        assert r <= TOL64
    # configurations differ (angles and grain draws), so the images must too
    assert _rel(imgs[0], imgs[1]) > 1e-3 and _rel(imgs[0], imgs[2]) > 1e-3


@pytest.mark.parametrize("name,flags,kw", [
    ("spheres_bcc", H.SMEAR_COMPAT, {"smear_mode": "compat"}),
    ("box_grating", H.SF_COMPAT_LB2, {"sf_flags": 1}),
    ("mesh_tetra", H.MESH_NO_TRANSLATION, {"mesh_translate": False}),
    ("cylinders_hcp", H.NO_SMEAR, {"smear_mode": "off"}),
])
def test_reference_compatible_switches(name, flags, kw):
    _say(name, "flags", flags)
    text = _text(name)
    if flags == H.SF_COMPAT_LB2:  # the lb[2] term needs a lattice vector b out of the plane
        text = text.replace("b = [0 1 0]", "b = [0 1 0.3]")
    img = H.HigSimulation(text=text, base_dir=str(GOLD), flags=flags).compute()
    ref = OH.run(text, base_dir=str(GOLD), **kw)
    r = _rel(img, ref)
    _say(name, "relL2", r)
    assert r <= TOL64
    # each switch changes the image against the default
    dflt = OH.run(text, base_dir=str(GOLD))
    assert _rel(ref, dflt) > 1e-6


@pytest.mark.parametrize("name", ["spheres_bcc", "cylinders_hcp", "mesh_tetra", "box_grating", "saxs_cubes"])
def test_fp32c_paths(name):
    _say(name, "fp32c")
    text = _text(name)
    sim = H.HigSimulation(text=text, base_dir=str(GOLD), precision="fp32c")
    img = sim.compute()
    ref = OH.run(text, base_dir=str(GOLD))
    r = _rel(img, ref)
    _say(name, "fp32c relL2", r)
    assert r <= TOL32
    if name == "mesh_tetra":
        # the mesh's form factors on the input's q-space grid (QGrid::create) run on the matrix cores
        paths = sim.last_paths()
        _say(name, "numeric form-factor paths", paths)
        assert paths["ewald"] > 0 and paths["pair"] == 0


def test_seed_changes_orientations_only_where_drawn():
    text = _text("spheres_bcc")
    a = H.HigSimulation(text=text, seed=1).compute()
    b = H.HigSimulation(text=text, seed=2).compute()
    assert _rel(a, b) > 1e-6
    ref = OH.run(text, seed=2)
    assert _rel(b, ref) <= TOL64
    # a single-orientation input draws nothing
    p1 = H.HigSimulation(text=_text("pyramid_dwba"), seed=1).compute()
    p2 = H.HigSimulation(text=_text("pyramid_dwba"), seed=99).compute()
    assert np.array_equal(p1, p2)


def _full_spheres(maxgrains):
    return _text("spheres_bcc").replace("resolution = [60 30]", "resolution = [600 300]").replace(
        "maxgrains = [4 4 1]", f"maxgrains = [{maxgrains} {maxgrains} 1]")


def test_full_detector_spheres_matches_oracle():
    """The detector of the reference's spheres example (600 x 300) with 25 grains, against the oracle."""
    text = _full_spheres(5)
    sim = H.HigSimulation(text=text)
    img = sim.compute()
    _say("full detector 25 grains", sim.last_counts())
    ref = OH.run(text)
    r = _rel(img, ref)
    _say("relL2", r)
    assert r <= TOL64


def test_full_spheres_example_paths_agree():
    """All 625 grains of the reference's spheres example at its 600 x 300 detector: the fused kernel and
    the per-grain calls agree (two independent device paths over the same sampled grains), and FP32C
    stays within its bound of FP64."""
    text = _full_spheres(25)
    fused = H.HigSimulation(text=text)
    a = fused.compute()
    ca = fused.last_counts()
    per = H.HigSimulation(text=text, flags=H.NO_FUSE)
    b = per.compute()
    cb = per.last_counts()
    _say("fused", ca, "per-grain", cb)
    assert ca["fused_grains"] == 625 and cb["grains"] == 625
    r = _rel(a, b)
    _say("fused vs per-grain relL2", r)
    assert r <= 1e-12
    c = H.HigSimulation(text=text, precision="fp32c").compute()
    r32 = _rel(c, a)
    _say("fp32c vs fp64 relL2", r32)
    assert r32 <= TOL32


def test_ensemble_kernel_matches_per_grain_calls_at_full_size():
    """All 625 grains of the reference's gratings example (boxes, 600 x 300) and the two-structure scan
    with rotated multi-element cells: the fused ensemble kernel against the per-grain calls."""
    box = (GOLD / "box_grating.hig").read_text().replace("resolution = [50 30]", "resolution = [600 300]").replace(
        "maxgrains = [3 3 1]", "maxgrains = [25 25 1]")
    a = H.HigSimulation(text=box)
    ia = a.compute()
    b = H.HigSimulation(text=box, flags=H.NO_FUSE)
    ib = b.compute()
    _say("gratings 625 grains: ensemble", a.last_counts(), "per grain", b.last_counts())
    assert a.last_counts()["fused_grains"] == 625
    r = _rel(ia, ib)
    _say("ensemble vs per-grain relL2", r)
    assert r <= 1e-12
    ic = H.HigSimulation(text=box, precision="fp32c").compute()
    r32 = _rel(ic, ia)
    _say("gratings fp32c vs fp64 relL2", r32)
    assert r32 <= TOL32
    text = _text("two_structures")
    fa = H.HigSimulation(text=text).run_all()
    fb = H.HigSimulation(text=text, flags=H.NO_FUSE).run_all()
    for x, y in zip(fa, fb):
        assert _rel(x, y) <= 1e-12


def _one_grain_pair(corr):
    t = _text("two_structures")
    t = t.replace("maxgrains = [2 2 1]", "maxgrains = [1 1 1]").replace("maxgrains = [2 1 1]", "maxgrains = [1 1 1]")
    return t.replace("resolution = [40 24]", f'resolution = [40 24], structcorrelation = "{corr}"')


@pytest.mark.parametrize("corr", ["GnE", "GE"])
def test_structure_correlations_match_oracle(corr):
    """Grain/ensemble-correlated structure sums (one grain per structure, two structures, smearing on)."""
    _say("structcorrelation", corr)
    text = _one_grain_pair(corr)
    img = H.HigSimulation(text=text).run_all()[0]  # the fixture scans angles: its first configuration
    ref = OH.run(text, config=0)
    r = _rel(img, ref)
    _say(corr, "relL2", r)
    assert r <= TOL64
    # the correlated sums differ from the plain one
    assert _rel(H.HigSimulation(text=_one_grain_pair("nGnE")).run_all()[0], ref) > 1e-2


def test_structure_correlations_refused_where_undefined():
    """nGE is unimplemented in the reference, and GnE / GE read past its per-structure array for more
    than one grain: both are refused with an error, not computed."""
    with pytest.raises(Exception, match="nGE"):
        H.HigSimulation(text=_one_grain_pair("nGE")).run_all()
    many = _text("two_structures").replace("resolution = [40 24]",
                                            'resolution = [40 24], structcorrelation = "GnE"')
    with pytest.raises(Exception, match="one grain"):
        H.HigSimulation(text=many).run_all()


def _spheres(grain):
    return _text("spheres_bcc").replace("repetition = [100 100 10], scaling = 15", grain)


_DISTS = {
    # all three axes distributed: samples i, i + n, i + 2n of the list are one step's scaling
    "scaling_abc": ('repetition = [20 20 4], scaling = { a = { stat = "gaussian", mean = 15, stddev = 0.3, '
                    'nsamples = 3 }, b = { stat = "gaussian", mean = 15, stddev = 0.2, nsamples = 2 }, '
                    'c = { stat = "cauchy", mean = 14, stddev = 0.25, nsamples = 3 } }'),
    # one axis distributed (given in the lattice block): the reference's indexing of its sample list
    "scaling_a_in_lattice": None,
    # repetition ranges: three repetition entries, grains by index modulo 3, the loop run three times
    "repetition_range": ('scaling = 15, repetitiondist = { xrepetition = { stat = "range", min = 5, max = 9 }, '
                         'yrepetition = { stat = "range", min = 4, max = 8 }, '
                         'zrepetition = { stat = "range", min = 2, max = 4 } }'),
}


def _dist_text(name):
    if name == "scaling_a_in_lattice":
        return _text("spheres_bcc").replace(
            'lattice = { type = "bcc" }',
            'lattice = { type = "bcc", a = { stat = "gaussian", mean = 15, stddev = 0.4, nsamples = 4 } }').replace(
            "repetition = [100 100 10], scaling = 15", "repetition = [20 20 4], scaling = [15 16 14]")
    return _spheres(_DISTS[name])


@pytest.mark.parametrize("flags", [0, H.NO_FUSE], ids=["fused", "per_grain"])
@pytest.mark.parametrize("name", list(_DISTS))
def test_grain_distributions_match_oracle(name, flags):
    """Grain scaling and repetition distributions (samples from the 64-bit Mersenne twister seeded 0)."""
    _say("distribution", name, "flags", flags)
    text = _dist_text(name)
    img = H.HigSimulation(text=text, flags=flags).compute()
    ref = OH.run(text)
    r = _rel(img, ref)
    _say(name, "relL2", r)
    assert r <= TOL64
    plain = OH.run(_spheres("repetition = [20 20 4], scaling = 15"))
    assert _rel(ref, plain) > 1e-3


def test_grain_distributions_refused_where_the_reference_fails():
    both = _spheres('scaling = { a = { stat = "gaussian", mean = 15, stddev = 0.3, nsamples = 3 }, '
                    'b = { stat = "gaussian", mean = 15, stddev = 0.3, nsamples = 3 }, '
                    'c = { stat = "gaussian", mean = 15, stddev = 0.3, nsamples = 3 } }, '
                    'repetitiondist = { xrepetition = { stat = "range", min = 5, max = 9 }, '
                    'yrepetition = { stat = "range", min = 4, max = 8 }, zrepetition = { stat = "range", min = 2, max = 3 } }')
    # nine scaling samples make nine repetition entries against three scaling steps
    with pytest.raises(Exception, match="not aligned"):
        H.HigSimulation(text=both).compute()
    gauss = _spheres('scaling = 15, repetitiondist = { xrepetition = { stat = "gaussian", min = 5, max = 9 }, '
                     'yrepetition = { stat = "range", min = 4, max = 8 }, zrepetition = { stat = "none", min = 2 } }')
    with pytest.raises(Exception, match="none or range"):
        H.HigSimulation(text=gauss).compute()


@pytest.mark.parametrize("flags", [0, H.NO_FUSE], ids=["fused", "per_grain"])
def test_truncated_spheres_match_oracle(flags):
    """Spheres with a height parameter: the reference's GPU path cuts them (driver flag), its CPU path
    ignores the height (default)."""
    text = _spheres("repetition = [20 20 4], scaling = 15").replace(
        'param = { type = "radius", min = 5 }',
        'param = { type = "radius", min = 5 }, param = { type = "height", min = 6, max = 8, stat = "uniform", nvalues = 3 }')
    cut = H.HigSimulation(text=text, flags=flags | H.TRUNCATED_SPHERE).compute()
    ref = OH.run(text, truncated_sphere=True)
    r = _rel(cut, ref)
    _say("truncated spheres relL2", r)
    assert r <= TOL64
    full = H.HigSimulation(text=text, flags=flags).compute()
    assert _rel(full, OH.run(text)) <= TOL64
    assert _rel(full, ref) > 1e-2


def test_errors_are_reported_not_fatal():
    text = _text("spheres_bcc").replace('unitcell:key = "cell"', 'unitcell:key = "nope"')
    sim = H.HigSimulation(text=text)
    with pytest.raises(Exception, match="nope"):
        sim.compute()
    # the driver stays usable after a failed run
    ok = H.HigSimulation(text=_text("spheres_bcc")).compute()
    assert np.all(np.isfinite(ok))


def test_run_all_writes_the_reference_files(tmp_path):
    """run_all_gisaxs's files: one .out (the image, 17 digits) and one .tif per configuration, named by
    the angles as the reference streams them."""
    text = _text("two_structures")
    sim = H.HigSimulation(text=text)
    paths = sim.run_all_gisaxs(str(tmp_path / "run"))
    names = sorted(Path(p).name for p in paths)
    _say("run_all files", names)
    assert len(names) == 8 and all(n.startswith(("img_ai=", "gisaxs_ai=")) for n in names)
    outs = sorted(p for p in paths if p.endswith(".out"))
    for c, p in enumerate(sorted(outs, key=lambda q: paths.index(q))):
        img = np.loadtxt(p)
        assert _rel(img, OH.run(text, config=c)) <= TOL64
