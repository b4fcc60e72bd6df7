"""Distance functors of the reference's fitting (include/analyzer/distance_functions.hpp) restated in
numpy over an observed image O, a model image I and a mask m (test infrastructure only).

norm_l2(x) = sqrt(sum m x^2) and vec_dot(a, b) = sum m |a b| with NaN terms dropped (:40-57); the
square-root distances sum (m |difference|)^2 with non-finite terms dropped, L2 (:125-143) keeps them."""
from __future__ import annotations

import numpy as np


def _nan0(x):
    return np.where(np.isnan(x), 0.0, x)


def norm_l2(x, m):
    return float(np.sqrt(np.sum(m * _nan0(x * x))))


def vec_dot(a, b, m):
    return float(np.sum(m * _nan0(np.abs(a * b))))


def _sumsq(t):
    t = np.where(np.isfinite(t), t, 0.0)
    return float(np.sum(t * t))


def chi2(O, I, m):
    """The build's default objective: sum m (I - O)^2 / max(O, 1)."""
    return float(np.sum(m * (I - O) ** 2 / np.maximum(O, 1.0)))


def l2(O, I, m):
    """AbsoluteDifferenceSquare (:125-143): every term kept, so a NaN pixel gives NaN."""
    t = m * np.abs(O - I)
    return float(np.sum(t * t))


def sqrt_unit_l2(O, I, m):
    """SqrtUnitVectorNormL2DistanceSquare (:423-456)."""
    with np.errstate(invalid="ignore", divide="ignore"):
        sr, sd = np.sqrt(O), np.sqrt(I)
        return _sumsq(m * np.abs(sr / norm_l2(sr, m) - sd / norm_l2(sd, m)))


def sqrt_c_l2(O, I, m):
    """SqrtCNormL2DistanceSquare (:460-499)."""
    with np.errstate(invalid="ignore", divide="ignore"):
        sr, sd = np.sqrt(O), np.sqrt(I)
        c = vec_dot(sd, sr, m) / norm_l2(sd, m) ** 2
        return _sumsq(m * np.abs(sr - c * sd))


def cbrt_unit_l2(O, I, m):
    """CbrtUnitVectorNormL2DistanceSquare (:543-575): the unit-vector form over cube roots, every term
    kept (the loop drops nothing)."""
    cr, cd = np.cbrt(O), np.cbrt(I)
    t = m * (cd / norm_l2(cd, m) - cr / norm_l2(cr, m))
    return float(np.sum(t * t))


def cbrt_c_l2(O, I, m):
    """CbrtCNormL2DistanceSquare (:579-616): the c-norm form over cube roots, every term kept."""
    cr, cd = np.cbrt(O), np.cbrt(I)
    c = vec_dot(cd, cr, m) / norm_l2(cd, m) ** 2
    t = m * np.abs(cr - c * cd)
    return float(np.sum(t * t))


METRICS = {"chi2": chi2, "l2": l2, "sqrt_unit_l2": sqrt_unit_l2, "sqrt_c_l2": sqrt_c_l2,
           "cbrt_unit_l2": cbrt_unit_l2, "cbrt_c_l2": cbrt_c_l2}
