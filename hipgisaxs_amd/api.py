"""Python face of the hgx C ABI: a Context owning a q-grid and meshes, with device outputs.

Device memory and the stream come from PyTorch (torch tensors on ``cuda``); every compute call
runs the HIP kernels of ``libhgx.so`` on the stream torch is currently using, so results can be
consumed by torch ops without extra synchronisation. Nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N


class HgxError(RuntimeError):
    pass


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One hgx context (device, stream, persistent q-grid and meshes)."""

    def __init__(self, device: int = 0):
        import torch

        if not torch.cuda.is_available():
            raise HgxError("no HIP device available")
        self._torch = torch
        self.lib = N.load()
        self.device = device
        h = C.c_void_p()
        rc = self.lib.hgx_create(device, C.byref(h))
        if rc != N.HGX_OK:
            raise HgxError(f"hgx_create failed: {N.ERRORS.get(rc, rc)}")
        self.h = h
        self.follow_torch_stream()

    # ----- plumbing ---------------------------------------------------------------------------
    def follow_torch_stream(self) -> None:
        s = self._torch.cuda.current_stream(self.device).cuda_stream
        self._check(self.lib.hgx_set_stream(self.h, C.c_void_p(s)), "set_stream")

    def _check(self, rc: int, what: str) -> None:
        if rc != N.HGX_OK:
            msg = self.lib.hgx_last_error(self.h)
            raise HgxError(f"{what}: {N.ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.hgx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        self._check(self.lib.hgx_synchronize(self.h), "synchronize")

    def profile_begin(self) -> None:
        self._check(self.lib.hgx_profile_begin(self.h), "profile_begin")

    def profile_end(self) -> dict:
        mm, am = C.c_double(), C.c_double()
        mn, an = C.c_int64(), C.c_int64()
        self._check(self.lib.hgx_profile_end(self.h, C.byref(mm), C.byref(mn), C.byref(am), C.byref(an)),
                    "profile_end")
        return {"main_ms": mm.value, "main_launches": mn.value, "aux_ms": am.value, "aux_launches": an.value}

    def _out(self, n: int, out=None):
        torch = self._torch
        if out is None:
            return torch.empty(n, dtype=torch.float64, device=f"cuda:{self.device}")
        if out.dtype != torch.float64 or out.numel() < n or not out.is_contiguous() or not out.is_cuda:
            raise HgxError("output must be a contiguous float64 cuda tensor of sufficient size")
        return out

    @staticmethod
    def as_complex(t):
        import torch

        return torch.view_as_complex(t.view(-1, 2))

    # ----- q-grid -----------------------------------------------------------------------------
    def qgrid_qspace(self, ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0):
        self._check(self.lib.hgx_qgrid_qspace(self.h, ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0),
                    "qgrid_qspace")

    def qgrid_explicit(self, ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0):
        self._check(self.lib.hgx_qgrid_explicit(self.h, ncol, nrow, qy_min, qz_min, qy_max, qz_max, alpha_i, k0),
                    "qgrid_explicit")

    def qgrid_detector(self, ncol, nrow, pixel, distance, col_center, row_horizon, alpha_i, k0):
        self._check(self.lib.hgx_qgrid_detector(self.h, ncol, nrow, pixel, distance, col_center, row_horizon,
                                                alpha_i, k0), "qgrid_detector")

    def qgrid_set(self, ncol, nrow, qx, qy, qz_ext, alpha_f=None):
        qx, qy = _f64(qx).ravel(), _f64(qy).ravel()
        qz = np.asarray(qz_ext)
        if np.iscomplexobj(qz):
            qz = np.stack([qz.real, qz.imag], -1)
        qz = _f64(qz).ravel()
        npix = ncol * nrow
        if qx.size != npix or qy.size != npix or qz.size not in (2 * npix, 8 * npix):
            raise HgxError("qgrid_set: array sizes do not match ncol*nrow")
        nb = qz.size // (2 * npix)
        af = None if alpha_f is None else _f64(alpha_f).ravel()
        self._check(self.lib.hgx_qgrid_set(self.h, ncol, nrow, nb, _ptr(qx), _ptr(qy), _ptr(qz), _ptr(af)),
                    "qgrid_set")

    def qz_extended(self, k0, alpha_i, dnl=0j):
        dnl = complex(dnl)
        self._check(self.lib.hgx_qz_extended(self.h, k0, alpha_i, dnl.real, dnl.imag), "qz_extended")

    def info(self):
        npix, nq = C.c_int64(), C.c_int64()
        nrow, ncol, nb = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.hgx_qgrid_info(self.h, C.byref(npix), C.byref(nq), C.byref(nrow), C.byref(ncol),
                                            C.byref(nb)), "qgrid_info")
        return {"npix": npix.value, "nq": nq.value, "nrow": nrow.value, "ncol": ncol.value, "nbranch": nb.value}

    def is_product(self) -> bool:
        v = C.c_int()
        self._check(self.lib.hgx_qgrid_is_product(self.h, C.byref(v)), "qgrid_is_product")
        return bool(v.value)

    def is_ewald(self):
        """(whether the grid has the Ewald-sphere form, the tile half-range of c times max |v| in 1/nm)"""
        v, dv = C.c_int(), C.c_double()
        self._check(self.lib.hgx_qgrid_is_ewald(self.h, C.byref(v), C.byref(dv)), "qgrid_is_ewald")
        return bool(v.value), dv.value

    def ff_numeric_path(self, mesh_id, method="approx", precision="fp64", flags=0):
        """("pair" | "product" | "ewald", series orders or -1): the kernels ff_numeric would run"""
        m = {"approx": N.FF_APPROX, "exact": N.FF_EXACT}[method]
        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        path, orders = C.c_int(), C.c_int()
        self._check(self.lib.hgx_ff_numeric_path(self.h, mesh_id, m, p, flags, C.byref(path), C.byref(orders)),
                    "ff_numeric_path")
        return {N.PATH_PAIR: "pair", N.PATH_PRODUCT: "product", N.PATH_EWALD: "ewald"}[path.value], orders.value

    def qgrid_get(self):
        i = self.info()
        qx = np.empty(i["npix"])
        qy = np.empty(i["npix"])
        qz = np.empty(2 * i["nq"])
        af = np.full(i["nrow"], np.nan)
        self._check(self.lib.hgx_qgrid_get(self.h, _ptr(qx), _ptr(qy), _ptr(qz), _ptr(af)), "qgrid_get")
        return qx, qy, qz[0::2] + 1j * qz[1::2], af

    # ----- meshes -----------------------------------------------------------------------------
    def upload_mesh(self, records, kind: str = "exact9") -> int:
        k = {"approx8": N.MESH_APPROX8, "exact9": N.MESH_EXACT9}[kind]
        width = 8 if k == N.MESH_APPROX8 else 9
        r = _f64(records).reshape(-1, width)
        mid = C.c_int()
        self._check(self.lib.hgx_mesh_upload(self.h, k, r.shape[0], _ptr(r), C.byref(mid)), "mesh_upload")
        if not hasattr(self, "_uploaded"):
            self._uploaded = {}
        self._uploaded[mid.value] = r.shape[0]
        return mid.value

    def release_mesh(self, mesh_id: int) -> None:
        self._check(self.lib.hgx_mesh_release(self.h, mesh_id), "mesh_release")

    def mesh_info(self, mesh_id: int):
        n, v, r = C.c_int64(), C.c_double(), C.c_double()
        self._check(self.lib.hgx_mesh_info(self.h, mesh_id, C.byref(n), C.byref(v), C.byref(r)), "mesh_info")
        return {"ntri": n.value, "volume": v.value, "radius": r.value}

    # ----- compute ----------------------------------------------------------------------------
    def ff_numeric(self, mesh_id, method="approx", precision="fp64", rot=None, transvec=None, flags=0, out=None):
        m = {"approx": N.FF_APPROX, "exact": N.FF_EXACT}[method]
        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        nq = self.info()["nq"]
        o = self._out(2 * nq, out)
        R = None if rot is None else _f64(rot).ravel()
        tv = None if transvec is None else _f64(transvec).ravel()
        self._check(self.lib.hgx_ff_numeric(self.h, mesh_id, m, p, flags, _ptr(R), _ptr(tv), C.c_void_p(o.data_ptr())),
                    "ff_numeric")
        return o

    def ff_numeric_adjoint(self, mesh_id, W, precision="fp64", rot=None, out=None):
        """Record adjoint of L = Re sum_q W(q) F(q) (F: approximate method): [ntri_uploaded, 6] device
        tensor of (dL/d(s n)_t, dL/dc_t). W: complex device tensor viewed as [2*nq] float64."""
        import torch

        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        nq = self.info()["nq"]
        if W.dtype != torch.float64 or W.numel() != 2 * nq or not W.is_cuda:
            raise HgxError("ff_numeric_adjoint: W must be a float64 device tensor of 2*nq values")
        n = self._uploaded[mesh_id]
        o = self._out(6 * n, out)
        R = None if rot is None else _f64(rot).ravel()
        self._check(self.lib.hgx_ff_numeric_adjoint(self.h, mesh_id, p, 0, _ptr(R), C.c_void_p(W.data_ptr()),
                                                    C.c_void_p(o.data_ptr())), "ff_numeric_adjoint")
        return o.view(n, 6)

    def ff_analytic(self, shape: str, dims: Sequence, precision="fp64", rot=None, transvec=None, tau=0.0, eta=0.0,
                    flags=0, out=None):
        """dims: per shape dimension a (values, weights) pair of 1-D arrays."""
        sp = N.ShapeParams()
        keep = []
        nd = 2 if shape == "sphere" and flags & N.SHAPE_TRUNCATED_SPHERE else N.SHAPE_DIMS[shape]
        if len(dims) != nd:
            raise HgxError(f"{shape} takes {nd} dimensions")
        for d, (v, w) in enumerate(dims):
            v, w = _f64(v).ravel(), _f64(w).ravel()
            keep += [v, w]
            sp.n[d] = v.size
            sp.val[d] = v.ctypes.data_as(C.POINTER(C.c_double))
            sp.wt[d] = w.ctypes.data_as(C.POINTER(C.c_double))
        sp.tau, sp.eta, sp.flags = tau, eta, flags
        nq = self.info()["nq"]
        o = self._out(2 * nq, out)
        R = None if rot is None else _f64(rot).ravel()
        tv = None if transvec is None else _f64(transvec).ravel()
        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        self._check(self.lib.hgx_ff_analytic(self.h, N.SHAPES[shape], C.byref(sp), p, _ptr(R), _ptr(tv),
                                             C.c_void_p(o.data_ptr())), "ff_analytic")
        return o

    def sf_lattice(self, la, lb, lc, repet, center=(0, 0, 0), dw_sigma=0.0, flags=0, rot=None, out=None):
        a, b, c, n, ce = (_f64(x).ravel() for x in (la, lb, lc, repet, center))
        nq = self.info()["nq"]
        o = self._out(2 * nq, out)
        R = None if rot is None else _f64(rot).ravel()
        self._check(self.lib.hgx_sf_lattice(self.h, _ptr(a), _ptr(b), _ptr(c), _ptr(n), _ptr(ce), dw_sigma, flags,
                                            _ptr(R), C.c_void_p(o.data_ptr())), "sf_lattice")
        return o

    def fresnel(self, delta, beta, thickness, order, k0, alpha_i, out=None):
        d, b, t = _f64(delta).ravel(), _f64(beta).ravel(), _f64(thickness).ravel()
        npix = self.info()["npix"]
        o = self._out(8 * npix, out)
        self._check(self.lib.hgx_fresnel(self.h, d.size, _ptr(d), _ptr(b), _ptr(t), order, k0, alpha_i,
                                         C.c_void_p(o.data_ptr())), "fresnel")
        return o

    def dwba_intensity(self, ff, fc=None, sf=None, weight=1.0, accumulate=False, out=None):
        npix = self.info()["npix"]
        o = self._out(npix, out)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        self._check(self.lib.hgx_dwba_intensity(self.h, ptr(fc), ptr(sf), ptr(ff), weight, int(accumulate),
                                                C.c_void_p(o.data_ptr())), "dwba_intensity")
        return o


def _fill_shape_params(sp, dims, keep):
    for d, (v, w) in enumerate(dims):
        v, w = _f64(v).ravel(), _f64(w).ravel()
        keep += [v, w]
        sp.n[d] = v.size
        sp.val[d] = v.ctypes.data_as(C.POINTER(C.c_double))
        sp.wt[d] = w.ctypes.data_as(C.POINTER(C.c_double))


def lattice_intensity(ctx: Context, shape: str, dims, la, lb, lc, repet, rots, weights=None, center=(0, 0, 0),
                      dw_sigma=0.0, dn2=1.0, sf_flags=0, precision="fp64", fc=None, accumulate=False, out=None):
    """hgx_lattice_intensity: I[p] (+)= sum_g |w_g dn2 sum_b fc_b F(R_g q_b) S(R_g q_b)|^2."""
    m = N.LatticeModel()
    keep = []
    m.shape = N.SHAPES[shape]
    _fill_shape_params(m.params, dims, keep)
    for name, v in (("la", la), ("lb", lb), ("lc", lc), ("repet", repet), ("center", center)):
        getattr(m, name)[:] = [float(x) for x in v]
    m.dw_sigma = dw_sigma
    m.dn2_re, m.dn2_im = complex(dn2).real, complex(dn2).imag
    m.sf_flags = sf_flags
    R = _f64(rots).reshape(-1, 9)
    keep.append(R)
    m.ngrains = R.shape[0]
    m.rot = R.ctypes.data_as(C.POINTER(C.c_double))
    if weights is not None:
        W = _f64(weights).ravel()
        keep.append(W)
        m.weight = W.ctypes.data_as(C.POINTER(C.c_double))
    npix = ctx.info()["npix"]
    o = ctx._out(npix, out)
    p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
    ctx._check(ctx.lib.hgx_lattice_intensity(ctx.h, C.byref(m), p, None if fc is None else C.c_void_p(fc.data_ptr()),
                                             int(accumulate), C.c_void_p(o.data_ptr())), "lattice_intensity")
    return o


def _ensemble_model(elements, la, lb, lc, repet, rots, centers, weights, dw_sigma, sf_flags, keep):
    el = (N.EnsembleElement * len(elements))()
    keep.append(el)
    for e, d in zip(el, elements):
        e.shape = N.SHAPES[d["shape"]]
        _fill_shape_params(e.params, d["dims"], keep)
        e.params.tau, e.params.eta = float(d.get("tau", 0.0)), float(d.get("eta", 0.0))
        e.rot[:] = [float(x) for x in np.asarray(d.get("rot", np.eye(3)), dtype=float).ravel()]
        L = _f64(d.get("locations", [[0.0, 0.0, 0.0]])).reshape(-1, 3)
        keep.append(L)
        e.nloc = L.shape[0]
        e.loc = L.ctypes.data_as(C.POINTER(C.c_double))
        e.dn2_re, e.dn2_im = complex(d.get("dn2", 1.0)).real, complex(d.get("dn2", 1.0)).imag
    m = N.EnsembleModel()
    m.nelem = len(elements)
    m.elem = el
    for name, v in (("la", la), ("lb", lb), ("lc", lc), ("repet", repet)):
        getattr(m, name)[:] = [float(x) for x in v]
    m.dw_sigma = dw_sigma
    m.sf_flags = sf_flags
    R = _f64(rots).reshape(-1, 9)
    keep.append(R)
    m.ngrains = R.shape[0]
    m.rot = R.ctypes.data_as(C.POINTER(C.c_double))
    for name, arr, w in (("center", centers, 3), ("weight", weights, 1)):
        if arr is not None:
            A = _f64(arr).ravel()
            keep.append(A)
            setattr(m, name, A.ctypes.data_as(C.POINTER(C.c_double)))
    return m


def ensemble_intensity(ctx: Context, elements, la, lb, lc, repet, rots, centers=None, weights=None, dw_sigma=0.0,
                       sf_flags=0, precision="fp64", fc=None, accumulate=False, out=None):
    """hgx_ensemble_intensity: a lattice of unit cells of analytic elements under several grains.
    elements: dicts with shape, dims [(values, weights), ...], rot (3x3, default identity), locations
    [[x, y, z], ...] and dn2 (complex)."""
    keep = []
    m = _ensemble_model(elements, la, lb, lc, repet, rots, centers, weights, dw_sigma, sf_flags, keep)
    o = ctx._out(ctx.info()["npix"], out)
    p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
    ctx._check(ctx.lib.hgx_ensemble_intensity(ctx.h, C.byref(m), p, None if fc is None else C.c_void_p(fc.data_ptr()),
                                              int(accumulate), C.c_void_p(o.data_ptr())), "ensemble_intensity")
    return o


def ensemble_intensity_tangent(ctx: Context, model: dict, tangents, fc=None):
    """hgx_ensemble_intensity_tangent: (I, dI [ndir][npix]) for a model given as the keyword arguments of
    ensemble_intensity (elements, la, lb, lc, repet, rots, centers, weights, dw_sigma, sf_flags) and a
    list of tangent dicts of the same form holding the derivatives of those fields (missing fields:
    zero; repet and shapes are taken from the model)."""
    keep = []
    base = dict(centers=None, weights=None, dw_sigma=0.0, sf_flags=0)
    base.update(model)
    m = _ensemble_model(base["elements"], base["la"], base["lb"], base["lc"], base["repet"], base["rots"],
                        base["centers"], base["weights"], base["dw_sigma"], base["sf_flags"], keep)
    ng = m.ngrains
    T = (N.EnsembleModel * len(tangents))()
    for k, t in enumerate(tangents):
        els = []
        for e, te in zip(base["elements"], t.get("elements", [{}] * len(base["elements"]))):
            dims = te.get("dims", [(np.zeros(len(v)), np.zeros(len(w))) for v, w in e["dims"]])
            els.append({"shape": e["shape"], "dims": dims, "tau": te.get("tau", 0.0), "eta": te.get("eta", 0.0),
                        "rot": te.get("rot", np.zeros((3, 3))),
                        "locations": te.get("locations", np.zeros_like(np.asarray(e.get("locations", [[0, 0, 0]]),
                                                                                  dtype=float))),
                        "dn2": te.get("dn2", 0.0)})
        T[k] = _ensemble_model(els, t.get("la", (0, 0, 0)), t.get("lb", (0, 0, 0)), t.get("lc", (0, 0, 0)),
                               base["repet"], t.get("rots", np.zeros((ng, 3, 3))), t.get("centers"),
                               t.get("weights"), t.get("dw_sigma", 0.0), base["sf_flags"], keep)
    npix = ctx.info()["npix"]
    I = ctx._out(npix, None)
    import torch

    dI = torch.empty(len(tangents) * npix, dtype=torch.float64, device=I.device)
    ctx._check(ctx.lib.hgx_ensemble_intensity_tangent(ctx.h, C.byref(m), len(tangents), T,
                                                      None if fc is None else C.c_void_p(fc.data_ptr()),
                                                      C.c_void_p(I.data_ptr()), C.c_void_p(dI.data_ptr())),
               "ensemble_intensity_tangent")
    return I, dI.view(len(tangents), npix)


def image_distance_tangent(ctx: Context, I, O, mask, dI, metric):
    """hgx_image_distance_tangent: (D, dD [ndir]) for device images I, O, optional mask and tangents dI
    [ndir][n]."""
    n = I.numel()
    ndir = dI.numel() // n
    D = C.c_double()
    dD = (C.c_double * ndir)()
    m = N.METRICS[metric] if isinstance(metric, str) else int(metric)
    ctx._check(ctx.lib.hgx_image_distance_tangent(ctx.h, n, m, C.c_void_p(I.data_ptr()), C.c_void_p(O.data_ptr()),
                                                  None if mask is None else C.c_void_p(mask.data_ptr()), ndir,
                                                  C.c_void_p(dI.data_ptr()), C.byref(D), dD), "image_distance_tangent")
    return D.value, np.array(list(dD))


def gaussian_smear(ctx: Context, image, ncol: int, nrow: int, sigma: float, compat: bool = False):
    """hgx_gaussian_smear in place on a device image [nrow * ncol]."""
    ctx._check(ctx.lib.hgx_gaussian_smear(ctx.h, int(ncol), int(nrow), float(sigma), N.SMEAR_COMPAT if compat else 0,
                                          C.c_void_p(image.data_ptr())), "gaussian_smear")
    return image


def image_distance(ctx: Context, I, O, mask=None, metric="sqrt_unit_l2") -> float:
    """hgx_image_distance of device images I and O (float64 tensors of one size), optional mask."""
    if I.numel() != O.numel() or (mask is not None and mask.numel() != I.numel()):
        raise ValueError("image_distance: I, O and mask must have one size")
    m = N.METRICS[metric] if isinstance(metric, str) else int(metric)
    out = C.c_double()
    ctx._check(ctx.lib.hgx_image_distance(ctx.h, I.numel(), m, C.c_void_p(I.data_ptr()), C.c_void_p(O.data_ptr()),
                                          None if mask is None else C.c_void_p(mask.data_ptr()), C.byref(out)),
               "image_distance")
    return out.value


class Batch:
    """hgx_batch_create / hgx_batch_run: items are (grid_context_or_None, mesh_id, rot or None, out tensor)."""

    def __init__(self, ctx: Context, items, precision="fp32c", flags=0):
        arr = (N.BatchItem * len(items))()
        self._keep = []
        for k, (gctx, mid, rot, out) in enumerate(items):
            arr[k].grid = None if gctx is None else gctx.h
            arr[k].mesh_id = mid
            arr[k].rot[:] = [float(x) for x in (np.eye(3) if rot is None else np.asarray(rot)).ravel()]
            arr[k].ff = C.c_void_p(out.data_ptr())
            self._keep.append(out)
        self.ctx = ctx
        bid = C.c_int()
        p = {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision]
        ctx._check(ctx.lib.hgx_batch_create(ctx.h, len(items), arr, p, flags, C.byref(bid)), "batch_create")
        self.id = bid.value

    def run(self, use_graph=False):
        self.ctx._check(self.ctx.lib.hgx_batch_run(self.ctx.h, self.id, int(use_graph)), "batch_run")

    def info(self):
        u, p = C.c_int64(), C.c_double()
        self.ctx._check(self.ctx.lib.hgx_batch_info(self.ctx.h, self.id, C.byref(u), C.byref(p)), "batch_info")
        return {"units": u.value, "pairs": p.value}

    def release(self):
        self.ctx._check(self.ctx.lib.hgx_batch_release(self.ctx.h, self.id), "batch_release")


def tune_batch(ctx: "Context", qpt: int, blocks_per_unit: int) -> None:
    """Work-unit shape of the context's later batches of per-pair kernels."""
    if N.load().hgx_tune_batch(ctx.h, qpt, blocks_per_unit) != N.HGX_OK:
        raise HgxError("tune_batch: invalid knob values")


# a new context's product-kernel setting (hgx_tune_product / hgx_tune_product_tile):
# (blocks per unit, real products per complex product, operands, row tiles, two-level, shared rows)
PRODUCT_DEFAULT = (64, 4, "f16x2", 2, True, True)


def tune_product(ctx: "Context", blocks_per_unit: int, complex_mults: int = 3, operands: str = "fp32",
                 row_tiles: int = 2, two_level: bool = True, shared_rows: bool = True) -> None:
    """Work-unit size (128-triangle blocks) of the product-grid matrix-core kernel, 3 or 4 real
    products per complex product, and the operand precision: "fp32" or "f16x2" (each operand split
    into two fp16 halves; 4 real products only). For "f16x2" also the 32-row tiles per wave (1, 2,
    4), two-level accumulation (False only with 4 row tiles) and whether the row operands are built
    once per workgroup and shared (2 or 4 row tiles)."""
    ops = {"fp32": N.PRODUCT_FP32, "f16x2": N.PRODUCT_F16X2}
    lib = N.load()
    if operands not in ops or lib.hgx_tune_product(ctx.h, blocks_per_unit, complex_mults,
                                                   ops[operands]) != N.HGX_OK:
        raise HgxError("tune_product: invalid knob values")
    if operands == "f16x2" and lib.hgx_tune_product_tile(ctx.h, int(row_tiles), int(bool(two_level)),
                                                         int(bool(shared_rows))) != N.HGX_OK:
        raise HgxError("tune_product: invalid row tiles / accumulation / sharing")


def tune_product64(ctx: "Context", complex_mults: int = 4) -> None:
    """Real products per complex product (4 or 3, Gauss) of the plain FP64 product-grid kernel."""
    if N.load().hgx_tune_product64(ctx.h, int(complex_mults)) != N.HGX_OK:
        raise HgxError("tune_product64: complex_mults must be 3 or 4")


def tune_adjoint(ctx: "Context", operands: str = "f16x2", col_tiles: int = 4) -> None:
    """Operands of the FP32C product-grid record adjoint: "f16x2" (split fp16, default; 4 (default)
    or 2 32-column tiles per wave) or "fp32"."""
    ops = {"fp32": N.PRODUCT_FP32, "f16x2": N.PRODUCT_F16X2}
    if operands not in ops or N.load().hgx_tune_adjoint(ctx.h, ops[operands], int(col_tiles)) != N.HGX_OK:
        raise HgxError("tune_adjoint: invalid operands / column tiles")


class FitPlan:
    """hgx_fit_plan_create / hgx_chi2_fwd / hgx_chi2_bwd. theta = (radius, height, a, tilt, sigma, scale)."""

    def __init__(self, ctx: Context, shape: str, ua, ub, uc, repet, I_obs, fc=None, mask=None, center=(0, 0, 0),
                 dn2=1.0, rot0=None, sf_flags=0):
        m = N.FitModel()
        m.shape = N.SHAPES[shape]
        for name, v in (("ua", ua), ("ub", ub), ("uc", uc), ("repet", repet), ("center", center)):
            getattr(m, name)[:] = [float(x) for x in v]
        m.dn2_re, m.dn2_im = complex(dn2).real, complex(dn2).imag
        m.rot0[:] = [float(x) for x in (np.eye(3) if rot0 is None else np.asarray(rot0)).ravel()]
        m.sf_flags = sf_flags
        self.ctx = ctx
        self._keep = (I_obs, fc, mask)  # the plan reads these device buffers by pointer
        pid = C.c_int()
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        ctx._check(ctx.lib.hgx_fit_plan_create(ctx.h, C.byref(m), ptr(fc), ptr(I_obs), ptr(mask), C.byref(pid)),
                   "fit_plan_create")
        self.id = pid.value

    def fwd(self, theta, precision="fp64", I_model=None) -> float:
        th = _f64(theta).ravel()
        chi2 = C.c_double()
        self.ctx._check(self.ctx.lib.hgx_chi2_fwd(self.ctx.h, self.id, th.ctypes.data_as(C.POINTER(C.c_double)),
                                                  {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision],
                                                  C.byref(chi2),
                                                  None if I_model is None else C.c_void_p(I_model.data_ptr())),
                        "chi2_fwd")
        return chi2.value

    def bwd(self, theta, precision="fp64"):
        th = _f64(theta).ravel()
        chi2 = C.c_double()
        g = np.zeros(6)
        self.ctx._check(self.ctx.lib.hgx_chi2_bwd(self.ctx.h, self.id, th.ctypes.data_as(C.POINTER(C.c_double)),
                                                  {"fp64": N.PREC_FP64, "fp32c": N.PREC_FP32C}[precision],
                                                  C.byref(chi2), g.ctypes.data_as(C.POINTER(C.c_double))),
                        "chi2_bwd")
        return chi2.value, g

    def metric(self, name: str) -> None:
        """Objective: "chi2" (default), "l2", "sqrt_unit_l2" or "sqrt_c_l2" (the reference's distance
        functors, include/analyzer/distance_functions.hpp)."""
        self.ctx._check(self.ctx.lib.hgx_fit_plan_metric(self.ctx.h, self.id, N.METRICS[name]), "fit_plan_metric")

    def fwd_many(self, thetas) -> np.ndarray:
        """Objective of each row of thetas [n, 6] (one synchronisation)."""
        th = _f64(thetas).reshape(-1, 6)
        out = np.zeros(th.shape[0])
        self.ctx._check(self.ctx.lib.hgx_chi2_fwd_many(self.ctx.h, self.id, th.shape[0],
                                                       th.ctypes.data_as(C.POINTER(C.c_double)), N.PREC_FP64,
                                                       out.ctypes.data_as(C.POINTER(C.c_double))), "chi2_fwd_many")
        return out

    def lbfgs(self, theta, free=(1, 1, 1, 1, 1, 1), lo=None, hi=None, iterations=100, tol=1e-9):
        """Bounded L-BFGS with the adjoint gradient; returns (theta, objective, iterations)."""
        th = _f64(theta).ravel().copy()
        fr = np.ascontiguousarray(free, dtype=np.int32)
        lo = _f64(np.full(6, -np.inf) if lo is None else lo).ravel()
        hi = _f64(np.full(6, np.inf) if hi is None else hi).ravel()
        f, it = C.c_double(), C.c_int()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.ctx._check(self.ctx.lib.hgx_fit_lbfgs(self.ctx.h, self.id, dp(th), fr.ctypes.data_as(C.POINTER(C.c_int)),
                                                   dp(lo), dp(hi), int(iterations), float(tol), C.byref(f),
                                                   C.byref(it)), "fit_lbfgs")
        return th, f.value, it.value

    def release(self):
        self.ctx._check(self.ctx.lib.hgx_fit_plan_release(self.ctx.h, self.id), "fit_plan_release")


def param_distribution(pmin, pmax, nvalues, stat=0, mean=float("nan"), sd=1.0, reference=False):
    """AnalyticFormFactor::param_distribution (src/ff/ff_ana.cpp:235-284) through the C ABI: exactly
    ``nvalues`` nodes with weights normalised to 1, or (``reference=True``) the reference's
    accumulating do-while with unnormalised weights."""
    lib = N.load()
    cap = max(4 * nvalues + 8, 16)
    v = np.empty(cap)
    w = np.empty(cap)
    n = lib.hgx_param_distribution(pmin, pmax, nvalues, stat, mean, sd, N.HGX_DIST_REFERENCE if reference else 0,
                                   v.ctypes.data_as(C.POINTER(C.c_double)),
                                   w.ctypes.data_as(C.POINTER(C.c_double)), cap)
    if n < 0:
        raise HgxError(f"param_distribution: {N.ERRORS.get(n, n)}")
    return v[:n].copy(), w[:n].copy()


def tune_numeric(ctx: "Context", qpt64: int, qpt32: int, target_wg: int) -> None:
    """Launch shape of the context's per-pair numeric kernels."""
    rc = N.load().hgx_tune_numeric(ctx.h, qpt64, qpt32, target_wg)
    if rc != N.HGX_OK:
        raise HgxError("tune_numeric: invalid knob values")


class GratingPlan:
    """hgx_grating_*: batched fit of trapezoidal grating lines. theta rows = (top width, sidewall angle in
    rad, height, scale)."""

    def __init__(self, ctx: Context, ncand: int, I_obs, fc=None, mask=None, length=100.0, pitch=100.0,
                 periods=200, n_bottom=64, n_side=40, n_top=64, n_length=256, dn2=1.0 + 0j):
        self.ctx, self.ncand = ctx, ncand
        m = N.GratingModel()
        m.length, m.pitch, m.periods = length, pitch, periods
        m.n_bottom, m.n_side, m.n_top, m.n_length = n_bottom, n_side, n_top, n_length
        m.dn2_re, m.dn2_im = complex(dn2).real, complex(dn2).imag
        self._keep = (I_obs, fc, mask)
        pid = C.c_int()
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
        ctx._check(ctx.lib.hgx_grating_plan_create(ctx.h, C.byref(m), ncand, ptr(fc), ptr(I_obs), ptr(mask),
                                                   C.byref(pid)), "grating_plan_create")
        self.id = pid.value

    def metric(self, name: str) -> None:
        """Objective: "chi2" (default), "l2", "sqrt_unit_l2" or "sqrt_c_l2"."""
        self.ctx._check(self.ctx.lib.hgx_grating_plan_metric(self.ctx.h, self.id, N.METRICS[name]),
                        "grating_plan_metric")

    def set_theta(self, theta):
        t = _f64(theta).reshape(self.ncand, 4)
        self.ctx._check(self.ctx.lib.hgx_grating_set_theta(self.ctx.h, self.id, _ptr(t)), "grating_set_theta")

    def eval(self, use_graph=False):
        self.ctx._check(self.ctx.lib.hgx_grating_eval(self.ctx.h, self.id, int(use_graph)), "grating_eval")

    def lbfgs(self, iterations: int):
        self.ctx._check(self.ctx.lib.hgx_grating_lbfgs(self.ctx.h, self.id, iterations), "grating_lbfgs")

    def get(self):
        th = np.empty((self.ncand, 4))
        c2 = np.empty(self.ncand)
        g = np.empty((self.ncand, 4))
        self.ctx._check(self.ctx.lib.hgx_grating_get(self.ctx.h, self.id, _ptr(th), _ptr(c2), _ptr(g)), "grating_get")
        return th, c2, g

    def intensity(self, cand: int, out=None):
        npix = self.ctx.info()["npix"]
        o = self.ctx._out(npix, out)
        self.ctx._check(self.ctx.lib.hgx_grating_intensity(self.ctx.h, self.id, cand, C.c_void_p(o.data_ptr())),
                        "grating_intensity")
        return o

    def release(self):
        if self.id is not None:
            self.ctx._check(self.ctx.lib.hgx_grating_plan_release(self.ctx.h, self.id), "grating_plan_release")
            self.id = None
