"""CPU oracle of the .hig simulation loop (test infrastructure only).

A restatement, in Python over the FP64 oracle kernels of this package, of what the reference does
with a .hig input: HiGInput's reading of the file (src/config/hig_input.cpp:62-2832, defaults from
src/model/{structure,shape,layer,inst_scattering,compute_params}.cpp), HipGISAXS::init
(src/sim/hipgisaxs_main.cpp:106-216), run_gisaxs (:692-1436) with spatial_distribution /
orientation_distribution (:1702-2086) and the smearing of include/numerics/convolutions.hpp:128-175.
Grain orientations drawn with rand() come from the C library's own rand() (the reference never
seeds it on this path, so a fresh process draws the seed-1 sequence). Only tests import this.

The fitting half (read_fitting, update_param, objective, MT19937_64, pso, bruteforce_points) restates
the analyzer. Pins: the generator by the C++ standard's 10000th-output value, the brute-force grid by
its float accumulation (0.1 added three times exceeds 0.3), the distance functors in oracle.metrics,
update_param by the grains it produces. Parity unpinned: the particle swarm's draw order and update
(the reference seeds it with the clock and holds no vectors), beyond convergence on a quadratic.
"""
from __future__ import annotations

import ctypes
import math
import re

import numpy as np

from . import (dwba_intensity, ff_analytic, ff_approx, ff_truncated_sphere, fresnel, param_distribution, qgrid_qspace,
               qz_extended, sf_lattice, tri9_to_rec8)

PI_ = 3.141592653589793
SQRT_2PI_ = 2.506628
DBL_MIN = 2.2250738585072014e-308

# ----- reading ---------------------------------------------------------------------------------

_TOK = re.compile(r"""\s*(?:(\#[^\n]*)|("[^"]*"|'[^']*')|([{}\[\]=,])|([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)|
                      ([A-Za-z_][A-Za-z0-9_:.\-]*))""", re.X)


def _tokens(text):
    pos, out = 0, []
    text = text.rstrip()
    while pos < len(text):
        m = _TOK.match(text, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot read .hig text at offset {pos}: {text[pos:pos + 20]!r}")
        pos = m.end()
        com, s, p, num, word = m.groups()
        if com is not None:
            continue
        if s is not None:
            out.append(("str", s[1:-1]))
        elif p is not None:
            out.append(("p", p))
        elif num is not None:
            out.append(("num", float(num)))
        elif word is not None:
            out.append(("str", word))
    return out


def parse_tree(text):
    """The file as nested lists of (key, value) pairs (objects), Python lists (arrays), floats and
    strings."""
    toks = _tokens(text)
    i = 0

    def value():
        nonlocal i
        kind, v = toks[i]
        if kind == "p" and v == "{":
            i += 1
            return assigns("}")
        if kind == "p" and v == "[":
            i += 1
            arr = []
            while toks[i] != ("p", "]"):
                arr.append(value())
                if toks[i] == ("p", ","):
                    i += 1
            i += 1
            return arr
        if kind in ("num", "str"):
            i += 1
            return v
        raise ValueError(f"unexpected token {v!r}")

    def assigns(close):
        nonlocal i
        out = []
        while True:
            if close is None and i >= len(toks):
                return out
            if close is not None and toks[i] == ("p", close):
                i += 1
                return out
            kind, key = toks[i]
            if kind != "str" or toks[i + 1] != ("p", "="):
                raise ValueError(f"expected 'key =' at {key!r}")
            i += 2
            out.append((key, value()))
            if i < len(toks) and toks[i] == ("p", ","):
                i += 1

    return assigns(None)


def _obj(pairs):
    return pairs if isinstance(pairs, list) and all(isinstance(p, tuple) for p in pairs) else None


def read_model(text):
    """The typed model the simulation reads, with the reference's defaults."""
    top = dict(parse_tree(text))["hipGisaxsInput"]
    m = {"shapes": {}, "unitcells": {}, "layers": {}, "structures": {},
         "scattering": {"expt": "gisaxs", "alphai": [0.1, 0.1, 0.1], "inplanerot": [0.0, 0.0, 0.0],
                        "tilt": [0.0, 0.0, 0.0], "photon": 10000.0, "unit": "ev", "smearing": 1.0},
         "compute": {"minpoint": [-1.0, -1.0], "maxpoint": [-1.0, -1.0], "resolution": [1, 1],
                     "region": "qspace", "nslices": 0, "structcorrelation": "nGnE"}}
    for key, val in top:
        if key == "shape":
            sh = {"xrot": 0.0, "yrot": 0.0, "zrot": 0.0, "delta": 0.0, "beta": 0.0, "params": {}}
            for k, v in val:
                if k in ("key", "name"):
                    sh[k] = v
                elif k in ("xrot", "yrot", "zrot"):
                    sh[k] = v * PI_ / 180.
                elif k == "refindex":
                    sh.update(dict(v))
                elif k == "param":
                    p = {"stat": "none", "min": 0.0, "max": 0.0, "p1": 0.0, "p2": 0.0, "nvalues": 1}
                    p.update(dict(v))
                    sh["params"][p["type"]] = p
            m["shapes"][sh["key"]] = sh
        elif key == "unitcell":
            d = dict(val)
            els = {}
            for k, v in val:
                if k == "element":
                    e = dict(v)
                    els[e["shape:key"]] = [list(map(float, loc)) for loc in e["locations"]]
            m["unitcells"][d["key"]] = els
        elif key == "layer":
            d = {"order": 0, "thickness": 0.0, "delta": 0.0, "beta": 0.0}
            for k, v in val:
                if k == "refindex":
                    d.update(dict(v))
                else:
                    d[k] = v
            m["layers"][int(d["order"])] = d
        elif key == "structure":
            st = {"iratio": 1.0, "lattice": "cubic", "hkl": "100", "abc": None, "repetition": [1.0, 1.0, 1.0],
                  "scaling": [1.0, 1.0, 1.0], "transvec": [0.0, 0.0, 0.0], "maxgrains": [1, 1, 1],
                  "scale_stat": ["none"] * 3, "scale_sd": [0.0] * 3, "scale_n": [40] * 3, "repdist": None,
                  "spacing": [0.0, 0.0, 0.0], "distribution": "regular", "stat": "single",
                  "rot": [{"axis": "x", "angles": [0.0, 0.0]}, {"axis": "y", "angles": [0.0, 0.0]},
                          {"axis": "z", "angles": [0.0, 0.0]}]}
            for k, v in val:
                if k in ("key", "iratio"):
                    st[k] = v
                elif k == "grain":
                    for gk, gv in v:
                        if gk == "unitcell:key":
                            st["unitcell"] = gv
                        elif gk == "lattice":
                            ld = dict(gv)
                            st["lattice"] = ld.get("type", "cubic")
                            st["hkl"] = str(ld.get("hkl", "100"))
                            for ax in "abc":  # a | b | c = { stat, mean, stddev, nsamples }: a scaling axis
                                if ax in ld and _is_obj(ld[ax]):
                                    _scaling_axis(st, "abc".index(ax), ld.pop(ax))
                            if "a" in ld:
                                st["abc"] = (ld["a"], ld["b"], ld["c"])
                        elif gk == "scaling":
                            if _is_obj(gv):
                                for ax, av in gv:
                                    _scaling_axis(st, "abc".index(ax), av)
                            else:
                                st["scaling"] = [gv] * 3 if isinstance(gv, float) else list(gv)
                        elif gk == "repetitiondist":
                            st["repdist"] = [{"stat": None, "min": 0.0, "max": 0.0} for _ in range(3)]
                            for rk, rv in gv:
                                d = st["repdist"]["xyz".index(rk[0])]
                                for qk, qv in rv:
                                    if qk in ("stat", "min", "max"):
                                        d[qk] = qv
                        elif gk in ("repetition", "transvec"):
                            st[gk] = list(gv)
                elif k == "ensemble":
                    for ek, ev in v:
                        if ek in ("maxgrains", "spacing"):
                            st[ek] = list(ev)
                        elif ek == "distribution":
                            st[ek] = ev
                        elif ek == "orientations":
                            for ok, ov in ev:
                                if ok == "stat":
                                    st["stat"] = ov
                                else:
                                    r = dict(ov)
                                    ang = list(r.get("angles", [0.0, 0.0]))
                                    st["rot"][int(ok[3]) - 1] = {
                                        "axis": r.get("axis", "x"), "angles": ang if len(ang) == 2 else ang * 2,
                                        "mean": r.get("anglemean", 0.0), "sd": r.get("anglesd", 0.0),
                                        "location": r.get("anglelocation", 0.0), "scale": r.get("anglescale", 0.0)}
            m["structures"][st["key"]] = st
        elif key == "instrumentation":
            for ik, iv in val:
                if ik != "scattering":
                    continue
                for sk, sv in iv:
                    sc = m["scattering"]
                    if sk in ("alphai", "inplanerot", "tilt"):
                        for rk, rv in sv:
                            sc[sk][{"min": 0, "max": 1, "step": 2}[rk]] = rv
                            if sk == "alphai" and rk == "step":
                                sc["inplanerot"][2] = rv  # the reference's switch falls through
                    elif sk == "photon":
                        d = dict(sv)
                        sc["photon"], sc["unit"] = d.get("value", 10000.0), d.get("unit", "ev")
                    elif sk in ("expt", "smearing"):
                        sc[sk] = sv
        elif key == "computation":
            for ck, cv in val:
                c = m["compute"]
                if ck == "outputregion":
                    for ok, ov in cv:
                        c["region" if ok == "type" else ok] = ov
                elif ck == "resolution":
                    c["resolution"] = [int(cv[0]), int(cv[1])]
                elif ck in ("nslices", "structcorrelation"):
                    c[ck] = cv
    return m


# ----- simulation ------------------------------------------------------------------------------

_LIBC = None


def _libc():
    global _LIBC
    if _LIBC is None:
        _LIBC = ctypes.CDLL("libc.so.6")
        _LIBC.rand.restype = ctypes.c_int
    return _LIBC


def libc_rand(seed):
    """The C library's rand() sequence after srand(seed), drawn on demand."""
    lib = _libc()
    lib.srand(ctypes.c_uint(seed))
    while True:
        yield lib.rand()


def rotmat(axis, a):
    c, s = math.cos(a), math.sin(a)
    if axis == 0:
        return np.array([[1., 0., 0.], [0., c, -s], [0., s, c]])
    if axis == 1:
        return np.array([[c, 0., -s], [0., 1., 0.], [s, 0., c]])
    return np.array([[c, -s, 0.], [s, c, 0.], [0., 0., 1.]])


def _mul(a, b):
    """Row-by-column product with the reference's summation order (k = 0, 1, 2)."""
    out = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            s = 0.0
            for k in range(3):
                s += a[i, k] * b[k, j]
            out[i, j] = s
    return out


def grains(st, phi, rand):
    """(rot, center, weight) per grain of a structure; rand() draws from the iterator `rand`."""
    nd = [int(x) for x in st["maxgrains"]]
    ng = nd[0] * nd[1] * nd[2]
    sp = st["spacing"]
    dx, dy, dz = [], [], []
    val = 0.0
    for i in range(nd[0]):
        dx.append(val - sp[0] * (nd[0] - 1) / 2)
        val += sp[0]
    val = 0.0
    for i in range(nd[1]):
        dy.append(val - sp[1] * (nd[1] - 1) / 2)
        val += sp[1]
    val = 0.0
    for i in range(nd[2]):
        dz.append(val)
        val += sp[2]
    rots = []
    for r in st["rot"]:
        if r["axis"] == "n":
            rots.append({"axis": "x", "angles": [0.0, 0.0], "mean": 0, "sd": 0, "location": 0, "scale": 0})
        else:
            rots.append(dict({"mean": 0, "sd": 0, "location": 0, "scale": 0}, **r))
    nn = np.zeros((3, ng))
    w = np.ones((3, ng))
    stat = st["stat"]
    weighted = False
    if stat == "single":
        for k in range(3):
            nn[k, :] = rots[k]["angles"][0] * PI_ / 180
    elif stat == "random":
        for k in range(3):
            for g in range(ng):
                nn[k, g] = (next(rand) / 2147483647.0) * 2 * PI_
    elif stat == "range":
        for k in range(3):
            lo, hi = rots[k]["angles"]
            d = abs(hi - lo)
            for g in range(ng):
                nn[k, g] = (lo + (next(rand) / 2147483647.0) * d) * PI_ / 180
    elif stat in ("gaussian", "normal", "cauchy"):
        # run_gisaxs weights the grains for "gaussian" and "cauchy" only (hipgisaxs_main.cpp:934-938)
        weighted = stat != "normal"
        for k in range(3):
            lo, hi = rots[k]["angles"]
            d = abs(hi - lo)
            if stat == "cauchy":
                d = d / (ng - 1) if ng > 1 else 0.0
            if d > DBL_MIN:
                if stat != "cauchy":
                    d /= ng
                for g in range(ng):
                    t = lo + g * d
                    nn[k, g] = t * PI_ / 180
                    if stat == "cauchy":
                        s, l = rots[k]["scale"], rots[k]["location"]
                        w[k, g] = (1.0 / (PI_ * s * (1.0 + ((t - l) / s) * ((t - l) / s))) if s > DBL_MIN
                                   else (0.0 if t - l > DBL_MIN else 1.0))
                    else:
                        sd, mean = rots[k]["sd"], rots[k]["mean"]
                        w[k, g] = ((1.0 if t - mean < DBL_MIN else 0.0) if sd < DBL_MIN else
                                   1.0 / (sd * SQRT_2PI_) * math.exp(-(t - mean) * (t - mean) / (2 * sd * sd)))
            else:
                nn[k, :] = lo * PI_ / 180
    else:
        raise ValueError(stat)
    out = []
    rphi = rotmat(2, phi)
    nyz = nd[1] * nd[2]
    for g in range(ng):
        R = rphi
        for k in (2, 1, 0):
            R = _mul(R, rotmat("xyz".index(rots[k]["axis"]), nn[k, g]))
        x, y, z = g // nyz, (g % nyz) // nd[2], g % nd[2]
        dd = np.array([dx[x], dy[y], dz[z]])
        c = np.array([R[i, 0] * dd[0] + R[i, 1] * dd[1] + R[i, 2] * dd[2] for i in range(3)])
        out.append((R, c + np.asarray(st["transvec"], float), float(w[0, g] * w[1, g] * w[2, g]) if weighted else 1.0))
    return out


_LATTICES = {
    ("cubic", "100"): ([1, 0, 0], [0, 1, 0], [0, 0, 1]),
    ("cubic", "110"): ([1, 0, 0], [0., 1.4142, 0], [0., 0.7071, 0.7071]),
    ("bcc", "100"): ([1., 0., 0.], [0., 1., 0.], [0.5, 0.5, 0.5]),
    ("bcc", "110"): ([1.0, 0.0, 0], [0.5, 0.7071, 0], [0.0, 0.7071, 0.7071]),
    ("fcc", "100"): ([0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]),
    ("fcc", "111"): ([1., 0., 0], [0.5, 0.866, -0.], [1., 0.5773, 0.866]),
    ("hcp", None): ([1, 0, 0], [0.5, 0.866, 0], [0.5, 0.2887, 0.8165]),
    ("hex", None): ([1, 0, 0], [0.5, math.sqrt(3.0) / 2.0, 0], [0, 0, 1]),
    ("fco", None): ([1, 0, 0], [0.5, 0.0, 0], [0, 0, 1.0]),
}

_DIMS = {"sphere": [["radius"]], "cylinder": [["radius"], ["height"]],
         "box": [["xsize", "edge"], ["ysize", "edge"], ["height", "edge"]], "cube": [["height"]],
         "prism3": [["edge"], ["height"]], "prism6": [["edge"], ["height"]],
         "pyramid": [["xsize"], ["ysize"], ["height"], ["baseangle"]]}
_STAT = {"none": 0, "single": 0, "uniform": 1, "gaussian": 2}


def shape_dims(sh):
    dims = []
    for names in _DIMS[sh["name"]]:
        p = next(sh["params"][n] for n in names if n in sh["params"])
        dims.append(param_distribution(p["min"], p["max"], int(p["nvalues"]), _STAT[p["stat"]], p["p1"], p["p2"],
                                       reference=True))
    return dims


def read_obj(path):
    """Triangles [n, 9] of a Wavefront OBJ file: faces split into fans (v0, vk, vk+1), indices from 1
    or negative from the last vertex read."""
    verts, tris = [], []
    for line in open(path):
        f = line.split()
        if not f:
            continue
        if f[0] == "v":
            verts.append([float(x) for x in f[1:4]])
        elif f[0] == "f":
            idx = []
            for t in f[1:]:
                i = int(t.split("/")[0])
                idx.append(i - 1 if i > 0 else len(verts) + i)
            for k in range(1, len(idx) - 1):
                tris.append(verts[idx[0]] + verts[idx[k]] + verts[idx[k + 1]])
    return np.array(tris, float)


def read_raw(path):
    """Approximate-method records [n, 8] of a raw shape file (7 numbers per triangle, pad 0)."""
    v = np.array(open(path).read().split(), float).reshape(-1, 7)
    return np.hstack([v, np.zeros((v.shape[0], 1))])


def mesh_ff(path, g, rot, loc, translate=True):
    """Approximate (centroid-rule) form factor of a mesh file shape, times exp(i (R q).t)."""
    rec = tri9_to_rec8(read_obj(path)) if path.endswith(".obj") else read_raw(path)
    F = ff_approx(rec, g, rot=rot)
    if translate:
        mq = g.q_rotated(rot)
        F = F * np.exp(1j * (mq[0] * loc[0] + mq[1] * loc[1] + mq[2] * loc[2]))
    return F


def smear(img, sigma, compat=False):
    """Convolutions::convolution_gaussian_2d on img [nrow][ncol] (compat: the reference's window and
    weight indexing; otherwise distance weights over the symmetric window)."""
    h = int(math.ceil(6 * sigma))
    g = np.array([1.0 / (sigma * math.sqrt(2 * PI_) * math.exp((x * x) / (2 * sigma * sigma)))
                  for x in range(-h, h + 1)])

    def one_d(line):
        n = line.size
        out = np.zeros(n)
        for i in range(n):
            beg = max(0, i - h)
            end = min(n, i + h) if compat else min(n - 1, i + h) + 1
            s = nrm = 0.0
            for k in range(beg, end):
                wk = g[k - beg] if compat else g[k - i + h]
                s += line[k] * wk
                nrm += wk
            out[i] = s / nrm if nrm > DBL_MIN else 0.0
        return out

    tmp = np.array([one_d(r) for r in img])
    return np.array([one_d(c) for c in tmp.T]).T


def _is_obj(v):
    return isinstance(v, list) and len(v) > 0 and all(isinstance(x, tuple) and len(x) == 2 for x in v)


def _scaling_axis(st, ax, fields):
    """One axis of a grain scaling distribution (GrainScaling, src/model/structure.cpp:35-46;
    src/config/hig_input.cpp:1206-1263, 1642-1652)."""
    for k, v in fields:
        if k == "stat":
            st["scale_stat"][ax] = "none" if v == "single" else v
        elif k == "mean":
            st["scaling"][ax] = v
        elif k == "stddev":
            st["scale_sd"][ax] = v
        elif k == "nsamples":
            st["scale_n"][ax] = int(v)


def grain_sets(st, gr):
    """run_gisaxs's grain scaling / repetition loop (src/sim/hipgisaxs_main.cpp:819-868, 1031-1061,
    2089-2226) for the grains gr [(R, center, w)]: a list of (scaling[3], repetition[3], grains with
    weights times the step's scaling weight), one per step and repetition class."""
    stat, sd, mean = st["scale_stat"], st["scale_sd"], st["scaling"]
    nv = [(1 if sd[i] == 0 else st["scale_n"][i]) if stat[i] != "none" else 1 for i in range(3)]
    maxn = max(nv)
    rng = MT19937_64(0)
    smp, wts = [], []
    for i in range(3):
        if stat[i] == "none":
            smp.append(mean[i])
            wts.append(1.0)
        elif stat[i] in ("gaussian", "cauchy"):
            lo, dx = mean[i] - 4 * sd[i], 8 * sd[i] / (maxn + 1)
            for j in range(maxn):
                t = lo + j * dx + rng.uniform() * dx
                smp.append(t)
                if stat[i] == "gaussian":  # HipGISAXS::gaussian (include/sim/hipgisaxs_main.hpp:151-156)
                    w = (1.0 if t - mean[i] < DBL_MIN else 0.0) if sd[i] < DBL_MIN else \
                        1.0 / (sd[i] * 2.506628) * math.exp(-(t - mean[i]) ** 2 / (2 * sd[i] * sd[i]))
                else:  # HipGISAXS::cauchy (:169-172)
                    w = 1.0 / (PI_ * sd[i] * (1.0 + ((t - mean[i]) / sd[i]) ** 2)) if sd[i] > DBL_MIN else \
                        (0.0 if t - mean[i] > DBL_MIN else 1.0)
                wts.append(w)
        else:
            raise ValueError(f"scaling statistic {stat[i]}")
    if st["repdist"] is not None:
        n = len(smp) if len(smp) > 1 else len(gr)
        vals = []
        for d in st["repdist"]:
            if d["stat"] in ("none", "single"):
                vals.append([float(d["min"])])
            elif d["stat"] == "range":
                r = MT19937_64(0)
                vals.append([float(math.floor(d["min"] + r.uniform() * (d["max"] - d["min"]))) for _ in range(n)])
            else:
                raise ValueError("repetition statistic")
        reps = [list(v) for v in zip(*vals)]
    else:
        reps = [list(st["repetition"])]
    nsc, nrep = len(smp) // 3, len(reps)
    if nsc == nrep or nrep == 1:
        nrs = nsc
    elif nsc == 1:
        nrs = nrep
    else:
        raise ValueError("scaling and repetition distributions are not aligned")
    isdist = any(x > 0 for x in sd)
    out = []
    for i in range(nrs):
        sc, w = ([smp[i], smp[i + nsc], smp[i + 2 * nsc]], wts[i]) if isdist else (smp[:3], wts[0])
        for r in range(nrep if st["repdist"] is not None else 1):
            gs = [(R, c, gw * w) for g, (R, c, gw) in enumerate(gr) if st["repdist"] is None or g % nrep == r]
            if gs:
                out.append((sc, reps[r], gs))
    return out


def run(text, seed=1, sf_flags=0, smear_mode="fixed", config=0, base_dir=".", mesh_translate=True, model=None,
        truncated_sphere=False):
    """Image [nrow][ncol] of configuration `config` (alpha_i outermost, as run_all_gisaxs loops) for
    .hig text (or an already read and edited `model`); grain orientations drawn from rand() after
    srand(seed) (earlier configurations draw first). truncated_sphere: spheres with a height parameter
    as the reference's GPU path cuts them (its CPU path ignores the height)."""
    m = read_model(text) if model is None else model
    sc, cp = m["scattering"], m["compute"]
    keV = sc["photon"] / 1000 if sc["unit"] == "ev" else sc["photon"]
    k0 = 2 * PI_ / (1.23984 / keV)
    ncol, nrow = cp["resolution"]
    npix = ncol * nrow
    a_min, a_max, a_step = sc["alphai"]
    a_max = max(a_max, a_min)
    na = 1 if (a_min == a_max or a_step == 0) else int((a_max - a_min) / a_step + 1)
    p_min, p_max, p_step = sc["inplanerot"]
    nph = 1 if p_step == 0 else int((p_max - p_min) / p_step + 1)
    t_min, t_max, t_step = sc["tilt"]
    nt = 1 if t_step == 0 else int((t_max - t_min) / t_step + 1)
    configs = []
    a = a_min
    for _ in range(na):
        p = p_min
        for _ in range(nph):
            for _ in range(nt):
                configs.append((a, p))
            p += p_step
        a += a_step
    draws = libc_rand(seed)
    ai_min = a_min * PI_ / 180
    base = qgrid_qspace(ncol, nrow, cp["minpoint"][0], cp["minpoint"][1], cp["maxpoint"][0], cp["maxpoint"][1],
                        ai_min, k0)
    gis = sc["expt"] == "gisaxs"
    for ci, (a_deg, p_deg) in enumerate(configs):
        alphai, phi = a_deg * PI_ / 180, p_deg * PI_ / 180
        g = qz_extended(base, k0, alphai, 0j) if gis else base
        nb = 4 if gis else 1
        layers = [m["layers"][o] for o in sorted(m["layers"]) if o >= 0]
        sub = m["layers"].get(-1, {"delta": 0.0, "beta": 0.0, "thickness": 0.0})
        stack = layers + [sub]
        fc = fresnel(g, [l["delta"] for l in stack], [l["beta"] for l in stack], [l["thickness"] for l in stack],
                     0, k0, alphai) if gis else None
        keys = sorted(m["structures"])
        ir = np.array([m["structures"][k]["iratio"] for k in keys], float)
        if ir.sum() != 1.0:
            ir = ir / ir.sum()
        corr = cp["structcorrelation"]
        if corr not in ("nGnE", "GnE", "GE"):
            raise ValueError(f"structcorrelation {corr}")
        total = np.zeros(npix)
        for si, key in enumerate(keys):
            st = m["structures"][key]
            gr = grains(st, phi, draws)
            if ci != config:
                continue
            # GnE / GE (src/sim/hipgisaxs_main.cpp:1231-1262, 1371-1398) read one grain-intensity
            # entry per grain from an array that holds one: defined for a single grain only
            if corr != "nGnE" and len(gr) != 1:
                raise ValueError("GnE/GE need one grain per structure")
            if st["abc"] is not None:
                va, vb, vc = st["abc"]
            else:
                va, vb, vc = _LATTICES.get((st["lattice"], st["hkl"]), _LATTICES.get((st["lattice"], None)))
            els = m["unitcells"][st["unitcell"]]
            I = np.zeros(npix)
            for s, repet, gset in grain_sets(st, gr):
                la, lb, lc = (np.asarray(va, float) * s[0], np.asarray(vb, float) * s[1],
                              np.asarray(vc, float) * s[2])
                for R, center, w in gset:
                    ff = np.zeros(g.nq, complex)
                    for skey in sorted(els):
                        sh = m["shapes"][skey]
                        dn2 = 0j - complex(2 * sh["delta"], 2 * sh["beta"])
                        srot = _mul(_mul(_mul(R, rotmat(0, sh["xrot"])), rotmat(1, sh["yrot"])),
                                    rotmat(2, sh["zrot"]))
                        for loc in els[skey]:
                            if sh["name"].endswith((".obj", ".dat")):
                                ff += dn2 * mesh_ff(f"{base_dir}/{sh['name']}", g, srot, loc, mesh_translate)
                            elif truncated_sphere and sh["name"] == "sphere" and "height" in sh["params"]:
                                hp = sh["params"]["height"]
                                hd = param_distribution(hp["min"], hp["max"], int(hp["nvalues"]), _STAT[hp["stat"]],
                                                        hp["p1"], hp["p2"], reference=True)
                                ff += dn2 * ff_truncated_sphere(shape_dims(sh) + [hd], g, rot=srot, transvec=loc)
                            else:
                                ff += dn2 * ff_analytic(sh["name"], shape_dims(sh), g, rot=srot, transvec=loc)
                    sf = sf_lattice(la, lb, lc, repet, center, g, 0.0, sf_flags, rot=R)
                    I = dwba_intensity(npix, nb, fc, sf, ff, w, I)
            # nGnE and GE sum iratio_s I_s (GE squares the sum below), GnE sums iratio_s I_s^2
            total += ir[si] * (I * I if corr == "GnE" else I)
        if ci == config:
            if corr == "GE":
                total = total * total
            img = total.reshape(nrow, ncol)
            if smear_mode != "off" and sc["smearing"] > DBL_MIN:
                img = smear(img, sc["smearing"], compat=smear_mode == "compat")
            return img
    raise IndexError(config)


# ----- fitting ---------------------------------------------------------------------------------
# The analyzer's side of a .hig input: the fitting block (src/config/hig_input.cpp:302-340,
# 849-890, 1444-1477, 1524-1571, 1755-1777), HiGInput::update_params (:2712-2832) over the model
# above, the objective of HipGISAXSObjectiveFunction (src/analyzer/objective_func_hipgisaxs.cpp:
# 245-332) and the particle-swarm and brute-force drivers (src/analyzer/hipgisaxs_fit_pso.cpp,
# hipgisaxs_fit_pso_particle.cpp, hipgisaxs_fit_bruteforce.cpp).


def read_fitting(text):
    """fitparams sorted by key as (key, variable, min, max, step, init), the reference data and the
    algorithm blocks in file order."""
    top = dict(parse_tree(text))["hipGisaxsInput"]
    params, ref, algos = {}, {}, []
    for key, val in top:
        if key != "fitting":
            continue
        for k, v in val:
            d = dict(v)
            if k == "fitparam":
                r = dict(d.get("range", []))
                params[d["key"]] = (d["key"], d["variable"], r.get("min", 0.0), r.get("max", 0.0),
                                    r.get("step", -1.0), d.get("init", 0.0))
            elif k == "referencedata":
                reg = dict(d.get("fitregion", []))
                ref = {"path": d.get("path", ""), "mask": d.get("mask", ""), "region": reg.get("type", "qspace"),
                       "regmin": reg.get("regmin"), "regmax": reg.get("regmax")}
            elif k == "algorithm":
                a = {"name": d.get("algoname", ""), "metric": d.get("distancemetric", ""),
                     "tolerance": d.get("tolerance", 0.0), "regularization": d.get("regularization", 0.0),
                     "params": {}}
                for ak, av in v:
                    if ak == "algoparam":
                        ad = dict(av)
                        a["params"][ad["type"]] = ad.get("pvalue", 0.0)
                algos.append(a)
    return [params[k] for k in sorted(params)], ref, algos


def _split(path):
    kw, _, rest = path.partition(":")
    mk = re.fullmatch(r"([^\[]*)\['(.*)'\]", kw)
    return (mk.group(1), mk.group(2), rest) if mk else (kw, "", rest)


_AX = {"x": 0, "0": 0, "y": 1, "1": 1, "z": 2, "2": 2}


def update_param(m, variable, v):
    """HiGInput::update_params for one path, on the model of read_model (Shape, Layer, Structure and
    ScatteringParams update_param: src/model/shape.cpp:145-247, layer.cpp:41-86, structure.cpp:519-820,
    inst_scattering.cpp:41-167). Raises KeyError on an unknown path."""
    name, key, rest = _split(variable)
    n1, key1, r1 = _split(rest)
    k2, _, r2 = r1.partition(":")
    k3, _, r3 = r2.partition(":")
    if name == "shape":
        sh = m["shapes"][key]
        if n1 in ("xrot", "yrot", "zrot"):
            sh[n1] = v  # stored as given (radians)
        elif n1 == "param":
            p = sh["params"][key1]
            if k2 not in ("min", "max", "p1", "p2", "nvalues"):
                raise KeyError(variable)
            p[k2] = int(v) if k2 == "nvalues" else v
        elif n1 != "originvec":  # the shape origin enters no simulation step
            raise KeyError(variable)
    elif name == "layer":
        lay = next(l for l in m["layers"].values() if l.get("key") == key)
        if n1 == "thickness":
            lay["thickness"] = v
        elif n1 == "refindex" and k2 in ("delta", "beta"):
            lay[k2] = v
        else:
            raise KeyError(variable)
    elif name == "structure":
        st = m["structures"][key]
        if n1 == "iratio":
            st["iratio"] = v
        elif n1 == "grain" and k2 == "lattice" and k3 in ("a", "b", "c"):
            if st["abc"] is None:
                vs = _LATTICES.get((st["lattice"], st["hkl"]), _LATTICES.get((st["lattice"], None)))
                st["abc"] = tuple(list(map(float, x)) for x in vs)
            else:
                st["abc"] = tuple(list(map(float, x)) for x in st["abc"])
            st["abc"]["abc".index(k3)][_AX[r3.partition(":")[0]]] = v
        elif n1 == "grain" and k2 == "scaling" and k3 == "":
            st["scaling"] = [v, v, v]
        elif n1 == "grain" and k2 == "scaling":
            st["scaling"]["abc".index(k3)] = v
        elif n1 == "grain" and k2 in ("transvec", "repetition"):
            st[k2][_AX[k3]] = v
        elif n1 == "ensemble" and k2 in ("spacing", "maxgrains"):
            st[k2][_AX[k3]] = v
        else:
            raise KeyError(variable)
    elif name == "instrumentation" and n1 == "scattering":
        sc = m["scattering"]
        if k2 in ("alphai", "inplanerot", "tilt"):
            sc[k2][{"min": 0, "max": 1, "step": 2}[k3]] = v
        elif k2 == "photon" and k3 == "value":
            sc["photon"] = v
        else:
            raise KeyError(variable)
    else:
        raise KeyError(variable)


def read_ascii_image(path):
    """ImageData::read (src/analyzer/ImageData.cpp:79-104): one row per non-empty line."""
    rows = [list(map(float, ln.split())) for ln in open(path) if ln.strip()]
    return np.array(rows, dtype=np.float64)


def read_edf_image(path):
    """EDFReader (src/file/edf_reader.cpp:24-150): `key = value ;` header items with white space
    dropped, closed by '}'; the data start at the 512-byte block after the one holding the '}' and are
    EDF_BinarySize bytes of little-endian float32, Dim_1 columns by Dim_2 rows (DataType and ByteOrder
    are not read)."""
    raw = open(path, "rb").read()
    white = b" \n\t\r"
    head, p, close = {}, 0, None
    while close is None:
        key, value = b"", b""
        while raw[p:p + 1] not in (b"=", b"}"):
            if raw[p] not in white:
                key += raw[p:p + 1]
            p += 1
        if raw[p:p + 1] == b"}":
            close = p
            break
        p += 1
        while raw[p:p + 1] != b";":
            if raw[p] not in white:
                value += raw[p:p + 1]
            p += 1
        p += 1
        head[key.decode()] = value.decode()
        q = p
        while raw[q] in white:
            q += 1
        if raw[q:q + 1] == b"}":
            close = q
    cols, rows, nb = (int(head.get(k, "0")) for k in ("Dim_1", "Dim_2", "EDF_BinarySize"))
    start = (close // 512 + 1) * 512
    assert nb // 4 == cols * rows, "mismatch in reference data size"
    return np.frombuffer(raw[start:start + nb], dtype="<f4").astype(np.float64).reshape(rows, cols)


def objective(text, params, x, metric, O, mask=None, reg=0.0, base_dir=".", seed=1):
    """HipGISAXSObjectiveFunction::operator(): the image of the edited model, its distance to O under
    `metric` (oracle.metrics), plus (reg / 2) |x - range midpoints|^2; non-finite -> 1e8."""
    from .metrics import METRICS
    m = read_model(text)
    pn = 0.0
    for p, xi in zip(params, x):
        update_param(m, p[1], xi)
        mean = (p[3] - p[2]) / 2.0 + p[2]
        pn += (xi - mean) ** 2
    img = run(None, model=m, base_dir=base_dir, seed=seed)
    mk = np.ones_like(O) if mask is None else mask
    d = METRICS[metric](O.ravel(), img.ravel(), mk.ravel())
    return d + reg / 2 * pn if math.isfinite(d) else 1e8


class MT19937_64:
    """std::mt19937_64 (the 64-bit Mersenne twister of C++11, [rand.predef]), the generator under
    woo::MTRandomNumberGenerator (include/woo/random/woo_mtrandom.hpp:25-33)."""
    _M = (1 << 64) - 1

    def __init__(self, seed=5489):
        mt = [seed & self._M]
        for i in range(1, 312):
            mt.append((6364136223846793005 * (mt[-1] ^ (mt[-1] >> 62)) + i) & self._M)
        self.mt, self.i = mt, 312

    def __call__(self):
        if self.i >= 312:
            mt = self.mt
            for k in range(312):
                y = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
                mt[k] = mt[(k + 156) % 312] ^ (y >> 1) ^ (0xB5026F5AA96619E9 if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self._M

    def uniform(self):
        """MTRandomNumberGenerator::rand: (g() - min) / (max - min) in double."""
        return float(self()) / float(self._M)


def pso(f, lo, hi, omega, phi1, phi2, nparticles, ngenerations, seed, tune_omega=False):
    """ParticleSwarmOptimization, base flavour, on objective f: returns (best x, best f, evaluated
    rows (x..., f) in order)."""
    d = len(lo)
    xmin = [min(a, b) for a, b in zip(lo, hi)]
    xmax = [max(a, b) for a, b in zip(lo, hi)]
    g = MT19937_64(seed)
    x, v = [], []
    for _ in range(nparticles):  # PSOParticle::init_random_uniform: value then velocity per parameter
        xi, vi = [], []
        for j in range(d):
            xi.append(xmin[j] + g.uniform() * (xmax[j] - xmin[j]))
            vi.append(-xmax[j] + g.uniform() * (xmax[j] + xmax[j]))
        x.append(xi)
        v.append(vi)
    pb, pf = [list(xi) for xi in x], [float("inf")] * nparticles
    best, fbest, rows = [0.0] * d, float("inf"), []
    for _ in range(ngenerations):
        for i in range(nparticles):
            fi = f(x[i])
            rows.append(list(x[i]) + [fi])
            if pf[i] > fi:
                pf[i], pb[i] = fi, list(x[i])
            if fbest > fi:
                fbest, best = fi, list(x[i])
        for i in range(nparticles):  # PSOParticle::update_particle
            for j in range(d):
                r1, r2 = g.uniform(), g.uniform()
                nv = omega * v[i][j] + phi1 * r1 * (pb[i][j] - x[i][j]) + phi2 * r2 * (best[j] - x[i][j])
                v[i][j] = min(max(nv, -xmax[j]), xmax[j])
                x[i][j] = min(max(x[i][j] + v[i][j], xmin[j]), xmax[j])
        if tune_omega:
            omega /= 2.0
    return best, fbest, rows


def bruteforce_points(lo, hi, step):
    """BruteForceOptimization::loop_over_param: p = min, min + step, ... <= max with the step
    accumulated (a negative step means 1), the first parameter outermost."""
    pts = [[]]
    for a, b, s in zip(lo, hi, step):
        s = 1.0 if s < 0 else s
        vals, p = [], a
        while p <= b:
            vals.append(p)
            p += s
        pts = [q + [val] for q in pts for val in vals]
    return pts
