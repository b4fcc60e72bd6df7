"""FP64 CPU oracle of the GISAXS path (test infrastructure only).

Loaded only by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py, as
the checker; the product library never uses it. The C++ sources are in ``hgx_oracle.cpp`` and
are built into ``oracle/liboracle.so`` by ``python -m hipgisaxs_amd.build oracle``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None
_vp, _i64, _d, _i = C.c_void_p, C.c_int64, C.c_double, C.c_int


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `python -m hipgisaxs_amd.build oracle`")
        lib = C.CDLL(str(LIB_PATH))
        sig = {
            "hgxo_qgrid_qspace": (_i, [_i, _i] + [_d] * 6 + [_vp] * 4),
            "hgxo_qgrid_explicit": (_i, [_i, _i] + [_d] * 6 + [_vp] * 4),
            "hgxo_qgrid_detector": (_i, [_i, _i] + [_d] * 6 + [_vp] * 4),
            "hgxo_qz_extended": (None, [_i64, _vp, _d, _d, _d, _d, _vp]),
            "hgxo_fresnel": (_i, [_i, _vp, _vp, _vp, _i, _d, _d, _i, _i, _vp, _vp]),
            "hgxo_ff_approx": (None, [_i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_exact": (None, [_i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i, _d, _vp]),
            "hgxo_ff_approx_adjoint": (None, [_i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_approx_as_written": (None, [_i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_param_distribution": (_i, [_d, _d, _i, _i, _d, _d, _i, _vp, _vp, _i]),
            "hgxo_ff_sphere": (None, [_i, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_cylinder": (None, [_i, _vp, _vp, _i, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_box": (None, [_i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp,
                                   _vp]),
            "hgxo_ff_cube": (None, [_i, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_prism3":(None, [_i, _vp, _vp, _i, _vp, _vp, _d, _d, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_prism6": (None, [_i, _vp, _vp, _i, _vp, _vp, _d, _d, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_ff_pyramid": (None, [_i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _i64, _i64, _vp,
                                       _vp, _vp, _vp, _vp, _vp]),
            "hgxo_sf_lattice": (None, [_vp, _vp, _vp, _vp, _vp, _d, _i, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
            "hgxo_dwba_intensity": (None, [_i64, _i, _vp, _vp, _vp, _d, _vp]),
            "hgxo_tri9_to_rec8": (_i64, [_i64, _vp, _vp]),
        }
        for k, (res, args) in sig.items():
            f = getattr(lib, k)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _a(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _p(x):
    return None if x is None else x.ctypes.data_as(C.c_void_p)


def _c2r(z) -> np.ndarray:
    z = np.asarray(z, dtype=np.complex128)
    return _a(np.stack([z.real, z.imag], -1).ravel())


def _r2c(x: np.ndarray) -> np.ndarray:
    return x[0::2] + 1j * x[1::2]


class Grid:
    """Host q-grid: qx, qy [npix], qz complex [nbranch*npix], alpha_f [nrow]."""

    def __init__(self, ncol, nrow, qx, qy, qz, alpha_f=None):
        self.ncol, self.nrow = ncol, nrow
        self.qx, self.qy = _a(qx).ravel(), _a(qy).ravel()
        self.qz = np.asarray(qz, dtype=np.complex128).ravel()
        self.alpha_f = None if alpha_f is None else _a(alpha_f)

    @property
    def npix(self):
        return self.ncol * self.nrow

    @property
    def nq(self):
        return self.qz.size

    def q_rotated(self, rot=None):
        R = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64).reshape(3, 3)
        p = np.arange(self.nq) % self.npix
        q = np.stack([self.qx[p] + 0j, self.qy[p] + 0j, self.qz])
        return R @ q


def _grid(fn, ncol, nrow, *args):
    npix = ncol * nrow
    qx, qy, qz, af = np.empty(npix), np.empty(npix), np.empty(npix), np.empty(nrow)
    rc = getattr(load(), fn)(ncol, nrow, *[float(a) for a in args], _p(qx), _p(qy), _p(qz), _p(af))
    if rc != 0:
        raise ValueError(f"{fn} rejected its arguments")
    return Grid(ncol, nrow, qx, qy, qz + 0j, af)


def qgrid_qspace(ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0) -> Grid:
    return _grid("hgxo_qgrid_qspace", ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0)


def qgrid_explicit(ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0) -> Grid:
    return _grid("hgxo_qgrid_explicit", ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0)


def qgrid_detector(ncol, nrow, pixel, distance, col_center, row_horizon, alpha_i, k0) -> Grid:
    return _grid("hgxo_qgrid_detector", ncol, nrow, pixel, distance, col_center, row_horizon, alpha_i, k0)


def qz_extended(g: Grid, k0, alpha_i, dnl=0j) -> Grid:
    out = np.empty(8 * g.npix)
    qz0 = _a(g.qz[: g.npix].real)
    load().hgxo_qz_extended(g.npix, _p(qz0), k0, alpha_i, complex(dnl).real, complex(dnl).imag, _p(out))
    return Grid(g.ncol, g.nrow, g.qx, g.qy, _r2c(out), g.alpha_f)


def fresnel(g: Grid, delta, beta, thickness, order, k0, alpha_i) -> np.ndarray:
    d, b, t = _a(delta), _a(beta), _a(thickness)
    fc = np.empty(8 * g.npix)
    rc = load().hgxo_fresnel(d.size, _p(d), _p(b), _p(t), order, k0, alpha_i, g.nrow, g.ncol, _p(g.alpha_f), _p(fc))
    if rc != 0:
        raise ValueError("fresnel rejected its arguments")
    return _r2c(fc)


def _qargs(g: Grid, rot):
    R = _a(np.eye(3) if rot is None else rot).ravel()
    return g.npix, g.nq, _p(g.qx), _p(g.qy), None, R


def ff_approx(rec8, g: Grid, rot=None) -> np.ndarray:
    r = _a(rec8).reshape(-1, 8)
    npix, nq, qx, qy, qz_arr, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    out = np.empty(2 * nq)
    load().hgxo_ff_approx(r.shape[0], _p(r), npix, nq, qx, qy, _p(qz), _p(R), _p(out))
    return _r2c(out)


def ff_approx_as_written(rec8, g: Grid, rot=None) -> np.ndarray:
    """ff_approx with the reference's per-pair rotation into a heap-allocated vector (timing baseline)."""
    r = _a(rec8).reshape(-1, 8)
    npix, nq, qx, qy, _, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    out = np.empty(2 * nq)
    load().hgxo_ff_approx_as_written(r.shape[0], _p(r), npix, nq, qx, qy, _p(qz), _p(R), _p(out))
    return _r2c(out)


def ff_approx_adjoint(rec8, g: Grid, W, rot=None) -> np.ndarray:
    """Record adjoint of L = Re sum_q W(q) F(q): [ntri, 6] = (dL/dA_t, dL/dc_t), A_t = s_t n_t."""
    r = _a(rec8).reshape(-1, 8)
    npix, nq, qx, qy, _, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    w = _c2r(np.asarray(W, dtype=np.complex128).ravel())
    out = np.empty(6 * r.shape[0])
    load().hgxo_ff_approx_adjoint(r.shape[0], _p(r), npix, nq, qx, qy, _p(qz), _p(R), _p(w), _p(out))
    return out.reshape(-1, 6)


def ff_exact(tri9, g: Grid, rot=None, sign=+1, tiny=0.0) -> np.ndarray:
    t = _a(tri9).reshape(-1, 9)
    npix, nq, qx, qy, _, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    out = np.empty(2 * nq)
    load().hgxo_ff_exact(t.shape[0], _p(t), npix, nq, qx, qy, _p(qz), _p(R), int(sign), float(tiny), _p(out))
    return _r2c(out)


def tri9_to_rec8(tri9) -> np.ndarray:
    t = _a(tri9).reshape(-1, 9)
    out = np.zeros((t.shape[0], 8))
    n = load().hgxo_tri9_to_rec8(t.shape[0], _p(t), _p(out))
    return out[:n].copy()


def param_distribution(pmin, pmax, nvalues, stat=0, mean=float("nan"), sd=1.0, reference=False):
    """Node values and weights of a shape parameter: exactly ``nvalues`` nodes with normalised
    weights, or (``reference=True``) the reference's accumulating do-while with unnormalised pdf."""
    cap = 4 * nvalues + 8
    v, w = np.empty(cap), np.empty(cap)
    n = load().hgxo_param_distribution(pmin, pmax, nvalues, stat, mean, sd, 1 if reference else 0, _p(v), _p(w),
                                       cap)
    if n < 0:
        raise ValueError("param_distribution rejected its arguments")
    return v[:n].copy(), w[:n].copy()


def _dims(dims):
    out = []
    for v, w in dims:
        v, w = _a(v).ravel(), _a(w).ravel()
        out.append((v.size, v, w))
    return out


def ff_truncated_sphere(dims, g: Grid, rot=None, transvec=None, slices=512) -> np.ndarray:
    """The reference GPU's sphere with heights (src/ff/gpu/ff_ana_sphere_gpu.cu:53-113): per radius R and
    height h, e^{i mq_z (h - R)} times the trapezoid (`slices` points, ends weighted 1/2) over z in
    [R - h, R] of the disks 2 pi r^2 J1(r Re q_par)/(r Re q_par) e^{i mq_z z}, r = sqrt(R^2 - z^2), a disk
    being 0 where |q_par| < 1e-13 or r Re q_par < 1e-13; summed with the node weights, times e^{i mq.t}."""
    from scipy.special import j1

    (rv, rw), (hv, hw) = ((np.asarray(v, float), np.asarray(w, float)) for v, w in dims)
    R = np.eye(3) if rot is None else np.asarray(rot, float).reshape(3, 3)
    pix = np.arange(g.nq) % g.npix
    q = np.stack([np.asarray(g.qx, float)[pix] + 0j, np.asarray(g.qy, float)[pix] + 0j, np.asarray(g.qz, complex)])
    mq = R @ q
    qp = np.sqrt(mq[0] * mq[0] + mq[1] * mq[1])
    ok = np.abs(qp) >= 1e-13
    out = np.zeros(g.nq, complex)
    for r0, wr in zip(rv, rw):
        for h0, wh in zip(hv, hw):
            dh = h0 / (slices - 1)
            acc = np.zeros(g.nq, complex)
            for s in range(slices):
                z = s * dh - (h0 - r0)
                rz = np.sqrt(max(r0 * r0 - z * z, 0.0))  # the reference: NaN past the pole by rounding
                t2 = rz * qp.real
                use = ok & (t2 >= 1e-13)
                disk = np.zeros(g.nq)
                disk[use] = 2 * np.pi * rz * rz * j1(t2[use]) / t2[use]
                acc += (1.0 if s in (0, slices - 1) else 2.0) * disk * np.exp(1j * mq[2] * z)
            out += wh * wr * np.exp(1j * mq[2] * (h0 - r0)) * (0.5 * dh) * acc
    if transvec is not None:
        t = np.asarray(transvec, float)
        out *= np.exp(1j * (mq[0] * t[0] + mq[1] * t[1] + mq[2] * t[2]))
    return out


def ff_analytic(shape: str, dims, g: Grid, rot=None, transvec=None, tau=0.0, eta=0.0, compat=0) -> np.ndarray:
    npix, nq, qx, qy, _, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    tv = None if transvec is None else _a(transvec)
    out = np.empty(2 * nq)
    D = _dims(dims)
    flat = []
    for n, v, w in D:
        flat += [n, _p(v), _p(w)]
    lib = load()
    tail = (npix, nq, qx, qy, _p(qz), _p(R), _p(tv), _p(out))
    if shape == "sphere":
        lib.hgxo_ff_sphere(*flat, *tail)
    elif shape == "cylinder":
        lib.hgxo_ff_cylinder(*flat, *tail)
    elif shape == "box":
        lib.hgxo_ff_box(*flat, *tail)
    elif shape == "cube":
        lib.hgxo_ff_cube(*flat, *tail)
    elif shape == "prism3":
        lib.hgxo_ff_prism3(*flat, tau, eta, *tail)
    elif shape == "prism6":
        lib.hgxo_ff_prism6(*flat, tau, eta, *tail)
    elif shape == "pyramid":
        lib.hgxo_ff_pyramid(*flat, int(compat), *tail)
    else:
        raise ValueError(shape)
    return _r2c(out)


def sf_lattice(la, lb, lc, repet, center, g: Grid, dw_sigma=0.0, flags=0, rot=None) -> np.ndarray:
    npix, nq, qx, qy, _, R = _qargs(g, rot)
    qz = _c2r(g.qz)
    a, b, c, n, ce = (_a(x) for x in (la, lb, lc, repet, center))
    out = np.empty(2 * nq)
    load().hgxo_sf_lattice(_p(a), _p(b), _p(c), _p(n), _p(ce), dw_sigma, flags, npix, nq, qx, qy, _p(qz), _p(R),
                           _p(out))
    return _r2c(out)


def dwba_intensity(npix, nbranch, fc, sf, ff, weight=1.0, I=None) -> np.ndarray:
    I = np.zeros(npix) if I is None else _a(I).copy()
    fca = None if fc is None else _c2r(fc)
    sfa = None if sf is None else _c2r(sf)
    ffa = _c2r(ff)
    load().hgxo_dwba_intensity(npix, nbranch, _p(fca), _p(sfa), _p(ffa), weight, _p(I))
    return I
