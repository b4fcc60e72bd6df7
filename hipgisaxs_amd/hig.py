"""Python binding of the .hig simulation driver (libhgx_hig.so, C ABI in include/hgx_hig.h).

``HigSimulation(path)`` reads a HipGISAXS input file and runs its scan of incidence and in-plane
rotation angles on the device, as ``hipgisaxs <input.hig>`` does (src/sim/hipgisaxs_sim.cpp:37-70);
images are numpy arrays [nrow, ncol] with row 0 at the highest q_z. There is no CPU fallback: a run
without a HIP device raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native as N
from .api import HgxError

LIB_PATH = Path(__file__).resolve().parent / "libhgx_hig.so"

NO_FUSE, SF_COMPAT_LB2, SMEAR_COMPAT, NO_SMEAR, MESH_NO_TRANSLATION, PROFILE = 1, 2, 4, 8, 16, 32
FIT_FD_GRADIENT, FIT_SEQUENTIAL, TRUNCATED_SPHERE = 64, 128, 256

_vp, _dp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
_SIGS = {
    "hgx_hig_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "hgx_hig_destroy": (C.c_int, [_vp]),
    "hgx_hig_load": (C.c_int, [_vp, C.c_char_p]),
    "hgx_hig_load_text": (C.c_int, [_vp, C.c_char_p, C.c_char_p]),
    "hgx_hig_options": (C.c_int, [_vp, C.c_int, C.c_int, C.c_uint]),
    "hgx_hig_init": (C.c_int, [_vp]),
    "hgx_hig_info": (C.c_int, [_vp, _ip, _ip, _ip, _dp]),
    "hgx_hig_config": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "hgx_hig_run": (C.c_int, [_vp, C.c_int, _vp]),
    "hgx_hig_compute": (C.c_int, [_vp, _vp]),
    "hgx_hig_grains": (C.c_int, [_vp, C.c_char_p, C.c_int, _vp, _vp, _vp, _ip]),
    "hgx_hig_last_counts": (C.c_int, [_vp, C.POINTER(C.c_int64), _dp]),
    "hgx_hig_kernel_ms": (C.c_int, [_vp, _dp, C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int64)]),
    "hgx_hig_last_paths": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "hgx_hig_save": (C.c_int, [_vp, C.c_int, C.c_int, C.c_char_p]),
    "hgx_hig_save_tiff": (C.c_int, [_vp, C.c_int, C.c_int, C.c_char_p]),
    "hgx_hig_read_mesh": (C.c_int, [C.c_char_p, _ip, C.POINTER(C.c_int64), C.c_int64, _vp]),
    "hgx_hig_update_param": (C.c_int, [_vp, C.c_char_p, C.c_double]),
    "hgx_hig_fit_info": (C.c_int, [_vp, _ip, _ip]),
    "hgx_hig_fit_param": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_size_t, _dp]),
    "hgx_hig_fit_algorithm": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_size_t, _ip, _dp]),
    "hgx_hig_fit_init": (C.c_int, [_vp]),
    "hgx_hig_fit_objective": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.c_double, _dp]),
    "hgx_hig_fit_gradient": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.c_double, _dp, _dp, C.POINTER(C.c_int)]),
    "hgx_hig_fit_run": (C.c_int, [_vp, C.c_int, C.c_ulonglong, _dp, _dp, C.POINTER(C.c_longlong)]),
    "hgx_hig_fit_history": (C.c_int, [_vp, C.c_longlong, _dp, C.POINTER(C.c_longlong)]),
    "hgx_hig_last_error": (C.c_char_p, [_vp]),
    "hgx_hig_convex_hull": (C.c_int64, [C.c_int64, _dp, C.c_int64, C.POINTER(C.c_int)]),
}
_lib = None


def load() -> C.CDLL:
    """Loads libhgx_hig.so after libhgx.so (and through it PyTorch's HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    N.load()
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hipgisaxs_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def read_mesh(path: str):
    """A custom shape's mesh file as (kind, records): "exact9" [ntri, 9] for .obj, "approx8" [ntri, 8]
    for .dat."""
    lib = load()
    kind, n = C.c_int(), C.c_int64()
    rc = lib.hgx_hig_read_mesh(str(path).encode(), C.byref(kind), C.byref(n), 0, None)
    if rc != 0:
        raise HgxError(f"hgx_hig_read_mesh: {N.ERRORS.get(rc, rc)}: cannot read {path}")
    w = 9 if kind.value == N.MESH_EXACT9 else 8
    rec = np.zeros((n.value, w))
    rc = lib.hgx_hig_read_mesh(str(path).encode(), C.byref(kind), C.byref(n), n.value, rec.ctypes.data_as(_vp))
    if rc != 0:
        raise HgxError(f"hgx_hig_read_mesh: {N.ERRORS.get(rc, rc)}")
    return ("exact9" if w == 9 else "approx8"), rec


def convex_hull(points) -> np.ndarray:
    """hgx_hig_convex_hull: triangles [ntri][3] (point indices, counter-clockwise seen from outside)
    of the convex hull of points [n][3]."""
    lib = load()
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    nt = lib.hgx_hig_convex_hull(p.shape[0], p.ctypes.data_as(_dp), 0, None)
    if nt < 0:
        raise HgxError(f"hgx_hig_convex_hull: {N.ERRORS.get(nt, nt)}")
    tri = np.zeros((nt, 3), dtype=np.int32)
    got = lib.hgx_hig_convex_hull(p.shape[0], p.ctypes.data_as(_dp), nt, tri.ctypes.data_as(C.POINTER(C.c_int)))
    if got != nt:
        raise HgxError("hgx_hig_convex_hull: inconsistent triangle count")
    return tri


class HigSimulation:
    """A .hig input bound to a device (HipGISAXS::construct_input + init)."""

    def __init__(self, path: str | None = None, text: str | None = None, base_dir: str = ".", device: int = 0,
                 precision: str = "fp64", flags: int = 0, seed: int = 1):
        self.lib = load()
        h = _vp()
        self._check(self.lib.hgx_hig_create(device, C.byref(h)), "create", h=None)
        self.h = h
        self.options(precision, flags, seed)
        if (path is None) == (text is None):
            raise ValueError("give exactly one of path and text")
        if path is not None:
            self._check(self.lib.hgx_hig_load(self.h, str(path).encode()), "load")
        else:
            self._check(self.lib.hgx_hig_load_text(self.h, text.encode(), str(base_dir).encode()), "load")

    def _check(self, rc, what, h=0):
        if rc != 0:
            msg = self.lib.hgx_hig_last_error(self.h if h == 0 else h)
            raise HgxError(f"hgx_hig_{what}: {N.ERRORS.get(rc, rc)}: {(msg or b'').decode()}")

    def options(self, precision="fp64", flags=0, seed=1):
        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        self._check(self.lib.hgx_hig_options(self.h, p, int(flags), int(seed)), "options")

    def close(self):
        if getattr(self, "h", None):
            self.lib.hgx_hig_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        nc, nr, ncfg, k0 = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.hgx_hig_info(self.h, C.byref(nc), C.byref(nr), C.byref(ncfg), C.byref(k0)), "info")
        return {"ncol": nc.value, "nrow": nr.value, "nconfig": ncfg.value, "k0": k0.value}

    def configurations(self) -> list[tuple[float, float, float]]:
        out = []
        for i in range(self.info()["nconfig"]):
            a, p, t = C.c_double(), C.c_double(), C.c_double()
            self._check(self.lib.hgx_hig_config(self.h, i, C.byref(a), C.byref(p), C.byref(t)), "config")
            out.append((a.value, p.value, t.value))
        return out

    def _image(self):
        i = self.info()
        return np.zeros((i["nrow"], i["ncol"]))

    def run(self, config: int = 0) -> np.ndarray:
        img = self._image()
        self._check(self.lib.hgx_hig_run(self.h, int(config), img.ctypes.data_as(_vp)), "run")
        return img

    def compute(self) -> np.ndarray:
        img = self._image()
        self._check(self.lib.hgx_hig_compute(self.h, img.ctypes.data_as(_vp)), "compute")
        return img

    def run_all(self) -> list[np.ndarray]:
        return [self.run(i) for i in range(self.info()["nconfig"])]

    def grains(self, key: str):
        n = C.c_int()
        self._check(self.lib.hgx_hig_grains(self.h, key.encode(), 0, None, None, None, C.byref(n)), "grains")
        rot, cen, w = np.zeros((n.value, 3, 3)), np.zeros((n.value, 3)), np.zeros(n.value)
        self._check(self.lib.hgx_hig_grains(self.h, key.encode(), n.value, rot.ctypes.data_as(_vp),
                                            cen.ctypes.data_as(_vp), w.ctypes.data_as(_vp), C.byref(n)), "grains")
        return rot, cen, w

    def last_counts(self) -> dict:
        c = (C.c_int64 * 4)()
        ms = C.c_double()
        self._check(self.lib.hgx_hig_last_counts(self.h, c, C.byref(ms)), "last_counts")
        return {"fused_grains": c[0], "grains": c[1], "form_factors": c[2], "structure_factors": c[3],
                "ms": ms.value}

    def last_paths(self) -> dict:
        """Numeric form-factor calls of the last run by kernel path."""
        c = (C.c_int64 * 3)()
        self._check(self.lib.hgx_hig_last_paths(self.h, c), "last_paths")
        return {"pair": c[0], "product": c[1], "ewald": c[2]}

    def kernel_ms(self) -> dict:
        """Kernel times of the last run (needs flags with PROFILE)."""
        m, a = C.c_double(), C.c_double()
        nm, na = C.c_int64(), C.c_int64()
        self._check(self.lib.hgx_hig_kernel_ms(self.h, C.byref(m), C.byref(nm), C.byref(a), C.byref(na)), "kernel_ms")
        return {"main_ms": m.value, "main_launches": nm.value, "aux_ms": a.value, "aux_launches": na.value}

    def run_all_gisaxs(self, out_dir: str) -> list[str]:
        """run_all_gisaxs with its files (src/sim/hipgisaxs_main.cpp:436-472): for every configuration
        img_ai=<a>_rot=<phi>_tilt=<t>.tif and gisaxs_ai=..._rot=..._tilt=....out, the angles printed as a
        C++ stream prints them (6 significant digits), into out_dir (created). Returns the written paths."""
        import os
        cfg = self.configurations()
        os.makedirs(out_dir, exist_ok=True)
        written = []
        for c, (a, p, t) in enumerate(cfg):
            img = self.run(c)
            tag = f"ai={a:g}_rot={p:g}_tilt={t:g}"
            tif, out = Path(out_dir) / f"img_{tag}.tif", Path(out_dir) / f"gisaxs_{tag}.out"
            self.save_tiff(img, str(tif))
            self.save(img, str(out))
            written += [str(tif), str(out)]
        return written

    # ---- fitting (the analyzer's HipGISAXSObjectiveFunction and fit drivers) ----------------------

    def update_param(self, variable: str, value: float) -> None:
        """HiGInput::update_params for one fit variable path (host only)."""
        self._check(self.lib.hgx_hig_update_param(self.h, variable.encode(), float(value)), "update_param")

    def fit_params(self) -> list[dict]:
        """The fitparam blocks in fit-vector order (by key): key, min, max, step (-1: none), init."""
        n, na = C.c_int(), C.c_int()
        self._check(self.lib.hgx_hig_fit_info(self.h, C.byref(n), C.byref(na)), "fit_info")
        out = []
        for i in range(n.value):
            key = C.create_string_buffer(256)
            r = (C.c_double * 4)()
            self._check(self.lib.hgx_hig_fit_param(self.h, i, key, 256, r), "fit_param")
            out.append({"key": key.value.decode(), "min": r[0], "max": r[1], "step": r[2], "init": r[3]})
        return out

    def fit_algorithms(self) -> list[dict]:
        """The algorithm blocks in file order: name, metric (HGX_METRIC_* id, -1 unsupported), tolerance,
        regularization."""
        n, na = C.c_int(), C.c_int()
        self._check(self.lib.hgx_hig_fit_info(self.h, C.byref(n), C.byref(na)), "fit_info")
        out = []
        for i in range(na.value):
            name = C.create_string_buffer(64)
            m = C.c_int()
            tr = (C.c_double * 2)()
            self._check(self.lib.hgx_hig_fit_algorithm(self.h, i, name, 64, C.byref(m), tr), "fit_algorithm")
            out.append({"name": name.value.decode(), "metric": m.value, "tolerance": tr[0], "regularization": tr[1]})
        return out

    def fit_init(self) -> None:
        self._check(self.lib.hgx_hig_fit_init(self.h), "fit_init")

    def objective(self, x, metric="sqrt_unit_l2", reg: float = 0.0) -> float:
        """HipGISAXSObjectiveFunction::operator(): distance of the image at fit vector x to the reference data."""
        xv = np.ascontiguousarray(x, dtype=np.float64)
        m = N.METRICS[metric] if isinstance(metric, str) else int(metric)
        v = C.c_double()
        self._check(self.lib.hgx_hig_fit_objective(self.h, xv.size, xv.ctypes.data_as(_dp), m, float(reg),
                                                   C.byref(v)), "fit_objective")
        return v.value

    def gradient(self, x, metric="sqrt_unit_l2", reg: float = 0.0):
        """The objective at x and its gradient in one tangent pass: (value, grad, computed), computed[j]
        False where the pass does not cover variable j (its grad[j] is 0)."""
        xv = np.ascontiguousarray(x, dtype=np.float64)
        m = N.METRICS[metric] if isinstance(metric, str) else int(metric)
        v = C.c_double()
        g = np.zeros(xv.size)
        done = (C.c_int * xv.size)()
        self._check(self.lib.hgx_hig_fit_gradient(self.h, xv.size, xv.ctypes.data_as(_dp), m, float(reg), C.byref(v),
                                                  g.ctypes.data_as(_dp), done), "fit_gradient")
        return v.value, g, np.array([bool(d) for d in done])

    def fit(self, algorithm: int = 0, seed: int = 0) -> dict:
        """Runs algorithm block `algorithm`; returns the best x, its objective and the evaluation count."""
        n = len(self.fit_params())
        x = np.zeros(n)
        v = C.c_double()
        ne = C.c_longlong()
        self._check(self.lib.hgx_hig_fit_run(self.h, int(algorithm), int(seed), x.ctypes.data_as(_dp), C.byref(v),
                                             C.byref(ne)), "fit_run")
        return {"x": x, "value": v.value, "evals": ne.value}

    def fit_history(self) -> np.ndarray:
        """Every objective evaluation since the last fit or fit_init: rows (x..., value)."""
        nr = C.c_longlong()
        self._check(self.lib.hgx_hig_fit_history(self.h, 0, None, C.byref(nr)), "fit_history")
        w = len(self.fit_params()) + 1
        rows = np.zeros((nr.value, w))
        self._check(self.lib.hgx_hig_fit_history(self.h, nr.value, rows.ctypes.data_as(_dp), C.byref(nr)),
                    "fit_history")
        return rows

    @staticmethod
    def save_tiff(image: np.ndarray, path: str) -> None:
        """The run's image file: log10 of the shifted image normalised to [0, 1], 16-bit grayscale TIFF."""
        lib = load()
        img = np.ascontiguousarray(image, dtype=np.float64)
        if lib.hgx_hig_save_tiff(img.ctypes.data_as(_vp), img.shape[1], img.shape[0], str(path).encode()) != 0:
            raise HgxError(f"hgx_hig_save_tiff: cannot write {path} (non-finite pixels or bad size)")

    @staticmethod
    def save(image: np.ndarray, path: str) -> None:
        lib = load()
        img = np.ascontiguousarray(image, dtype=np.float64)
        if lib.hgx_hig_save(img.ctypes.data_as(_vp), img.shape[1], img.shape[0], str(path).encode()) != 0:
            raise HgxError(f"hgx_hig_save: cannot write {path}")
