"""ctypes declarations of the hgx C ABI (include/hgx.h) and a loader that fails loudly.

The shared library is ``hipgisaxs_amd/libhgx.so`` built by ``python -m hipgisaxs_amd.build``.
There is no fallback: if the library is missing or a HIP device is absent the calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhgx.so"

HGX_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "out of memory",
    -3: "device error",
    -4: "no q-grid",
    -5: "unsupported",
    -6: "bad handle",
}

PREC_FP64, PREC_FP32C = 0, 1
MESH_APPROX8, MESH_EXACT9 = 0, 1
FF_APPROX, FF_EXACT = 0, 1
FLAG_EXACT_REFERENCE_SIGN, FLAG_ACC_KAHAN, FLAG_NO_MFMA, FLAG_EXACT_CPU_THRESHOLD = 1, 2, 4, 8
PRODUCT_FP32, PRODUCT_F16X2 = 0, 1
PATH_PAIR, PATH_PRODUCT, PATH_EWALD = 0, 1, 2
SF_COMPAT_LB2 = 1
SF_FP32C = 2
SMEAR_COMPAT = 1
SHAPE_COMPAT_PYRAMID, SHAPE_TRUNCATED_SPHERE = 1, 2
HGX_DIST_REFERENCE = 1
METRICS = {"chi2": 0, "l2": 1, "sqrt_unit_l2": 2, "sqrt_c_l2": 3, "cbrt_unit_l2": 4, "cbrt_c_l2": 5}
SHAPES = {"sphere": 0, "cylinder": 1, "box": 2, "cube": 3, "prism3": 4, "prism6": 5, "pyramid": 6}
SHAPE_DIMS = {"sphere": 1, "cylinder": 2, "box": 3, "cube": 1, "prism3": 2, "prism6": 2, "pyramid": 4}
MAX_NODES = 40

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class ShapeParams(C.Structure):
    _fields_ = [
        ("n", C.c_int * 4),
        ("val", _dp * 4),
        ("wt", _dp * 4),
        ("tau", C.c_double),
        ("eta", C.c_double),
        ("flags", C.c_int),
    ]


class LatticeModel(C.Structure):
    _fields_ = [
        ("shape", C.c_int),
        ("params", ShapeParams),
        ("la", C.c_double * 3),
        ("lb", C.c_double * 3),
        ("lc", C.c_double * 3),
        ("repet", C.c_double * 3),
        ("center", C.c_double * 3),
        ("dw_sigma", C.c_double),
        ("dn2_re", C.c_double),
        ("dn2_im", C.c_double),
        ("sf_flags", C.c_int),
        ("ngrains", C.c_int),
        ("rot", _dp),
        ("weight", _dp),
    ]


class EnsembleElement(C.Structure):
    _fields_ = [
        ("shape", C.c_int),
        ("params", ShapeParams),
        ("rot", C.c_double * 9),
        ("nloc", C.c_int),
        ("loc", _dp),
        ("dn2_re", C.c_double),
        ("dn2_im", C.c_double),
    ]


class EnsembleModel(C.Structure):
    _fields_ = [
        ("nelem", C.c_int),
        ("elem", C.POINTER(EnsembleElement)),
        ("la", C.c_double * 3),
        ("lb", C.c_double * 3),
        ("lc", C.c_double * 3),
        ("repet", C.c_double * 3),
        ("dw_sigma", C.c_double),
        ("sf_flags", C.c_int),
        ("ngrains", C.c_int),
        ("rot", _dp),
        ("center", _dp),
        ("weight", _dp),
    ]


class FitModel(C.Structure):
    _fields_ = [
        ("shape", C.c_int),
        ("ua", C.c_double * 3),
        ("ub", C.c_double * 3),
        ("uc", C.c_double * 3),
        ("repet", C.c_double * 3),
        ("center", C.c_double * 3),
        ("dn2_re", C.c_double),
        ("dn2_im", C.c_double),
        ("rot0", C.c_double * 9),
        ("sf_flags", C.c_int),
    ]


class GratingModel(C.Structure):
    _fields_ = [
        ("length", C.c_double),
        ("pitch", C.c_double),
        ("periods", C.c_int),
        ("n_bottom", C.c_int),
        ("n_side", C.c_int),
        ("n_top", C.c_int),
        ("n_length", C.c_int),
        ("dn2_re", C.c_double),
        ("dn2_im", C.c_double),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("grid", _vp),
        ("mesh_id", C.c_int),
        ("rot", C.c_double * 9),
        ("ff", _vp),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "hgx_version": (C.c_char_p, []),
    "hgx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "hgx_destroy": (C.c_int, [_vp]),
    "hgx_set_stream": (C.c_int, [_vp, _vp]),
    "hgx_use_own_stream": (C.c_int, [_vp]),
    "hgx_synchronize": (C.c_int, [_vp]),
    "hgx_last_error": (C.c_char_p, [_vp]),
    "hgx_device_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "hgx_device_free": (C.c_int, [_vp, _vp]),
    "hgx_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "hgx_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "hgx_profile_begin": (C.c_int, [_vp]),
    "hgx_profile_end": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64)]),
    "hgx_qgrid_qspace": (C.c_int, [_vp, C.c_int, C.c_int] + [C.c_double] * 6),
    "hgx_qgrid_explicit": (C.c_int, [_vp, C.c_int, C.c_int] + [C.c_double] * 6),
    "hgx_qgrid_detector": (C.c_int, [_vp, C.c_int, C.c_int] + [C.c_double] * 6),
    "hgx_qgrid_set": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "hgx_qz_extended": (C.c_int, [_vp] + [C.c_double] * 4),
    "hgx_qgrid_info": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hgx_qgrid_is_product": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "hgx_qgrid_is_ewald": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "hgx_tune_ewald": (C.c_int, [_vp, C.c_int]),
    "hgx_ff_numeric_path": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "hgx_qgrid_get": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "hgx_mesh_upload": (C.c_int, [_vp, C.c_int, C.c_int64, _vp, C.POINTER(C.c_int)]),
    "hgx_mesh_release": (C.c_int, [_vp, C.c_int]),
    "hgx_mesh_info": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hgx_ff_numeric": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hgx_ff_numeric_adjoint": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hgx_ff_analytic": (C.c_int, [_vp, C.c_int, C.POINTER(ShapeParams), C.c_int, _vp, _vp, _vp]),
    "hgx_sf_lattice": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp]),
    "hgx_fresnel": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, _vp]),
    "hgx_dwba_intensity": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp]),
    "hgx_lattice_intensity": (C.c_int, [_vp, C.POINTER(LatticeModel), C.c_int, _vp, C.c_int, _vp]),
    "hgx_ensemble_intensity": (C.c_int, [_vp, C.POINTER(EnsembleModel), C.c_int, _vp, C.c_int, _vp]),
    "hgx_ensemble_intensity_tangent": (C.c_int, [_vp, C.POINTER(EnsembleModel), C.c_int, C.POINTER(EnsembleModel),
                                                 _vp, _vp, _vp]),
    "hgx_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(BatchItem), C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "hgx_batch_run": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_batch_info": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "hgx_batch_release": (C.c_int, [_vp, C.c_int]),
    "hgx_ff_numeric_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hgx_tune_batch": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_fit_plan_create": (C.c_int, [_vp, C.POINTER(FitModel), _vp, _vp, _vp, C.POINTER(C.c_int)]),
    "hgx_fit_plan_release": (C.c_int, [_vp, C.c_int]),
    "hgx_chi2_fwd": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.POINTER(C.c_double), _vp]),
    "hgx_chi2_bwd": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.POINTER(C.c_double), _dp]),
    "hgx_fit_plan_metric": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_chi2_fwd_many": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_int, _dp]),
    "hgx_fit_lbfgs": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int), _dp, _dp, C.c_int, C.c_double, _dp,
                                C.POINTER(C.c_int)]),
    "hgx_param_distribution": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_int, _dp, _dp, C.c_int]),
    "hgx_tune_numeric": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "hgx_tune_product": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "hgx_tune_product64": (C.c_int, [_vp, C.c_int]),
    "hgx_tune_product_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "hgx_tune_adjoint": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_caxpy": (C.c_int, [_vp, C.c_int64, C.c_double, C.c_double, _vp, C.c_int, _vp]),
    "hgx_axpy": (C.c_int, [_vp, C.c_int64, C.c_double, _vp, C.c_int, _vp]),
    "hgx_axpy_sq": (C.c_int, [_vp, C.c_int64, C.c_double, _vp, C.c_int, _vp]),
    "hgx_gaussian_smear": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, _vp]),
    "hgx_image_distance": (C.c_int, [_vp, C.c_int64, C.c_int, _vp, _vp, _vp, C.POINTER(C.c_double)]),
    "hgx_image_distance_async": (C.c_int, [_vp, C.c_int64, C.c_int, _vp, _vp, _vp, _vp]),
    "hgx_image_distance_tangent": (C.c_int, [_vp, C.c_int64, C.c_int, _vp, _vp, _vp, C.c_int, _vp,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hgx_grating_plan_create": (C.c_int, [_vp, C.POINTER(GratingModel), C.c_int, _vp, _vp, _vp, C.POINTER(C.c_int)]),
    "hgx_grating_plan_release": (C.c_int, [_vp, C.c_int]),
    "hgx_grating_plan_metric": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_grating_set_theta": (C.c_int, [_vp, C.c_int, _vp]),
    "hgx_grating_eval": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_grating_lbfgs": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hgx_grating_get": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "hgx_grating_intensity": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libhgx.so (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hipgisaxs_amd.build`")
    # PyTorch ships its own HIP runtime; it must be the process's only one, so it is loaded before
    # libhgx.so, whose libamdhip64.so.7 dependency then binds to it (loading libhgx.so first would
    # bring in a second runtime and the device calls of both would fail)
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)
