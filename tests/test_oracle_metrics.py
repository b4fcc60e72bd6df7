"""Pins of the oracle's restatement of the reference's distance functors
(include/analyzer/distance_functions.hpp): closed forms and invariances that a dropped square root,
a wrong normalisation or a mask applied once instead of twice would break."""
import numpy as np
import pytest

from oracle import metrics as M


@pytest.fixture
def images():
    rng = np.random.default_rng(11)
    I = rng.gamma(2.0, 3.0, 500)
    O = rng.poisson(I).astype(float)
    m = (rng.uniform(size=500) > 0.1).astype(float)
    return O, I, m


def test_scaled_data_are_at_distance_zero(images):
    """sqrt_unit_l2 and sqrt_c_l2 are invariant to the scale of either image: O = k I gives 0."""
    _, I, m = images
    for k in (1.0, 3.7, 1e-4):
        assert M.sqrt_unit_l2(k * I, I, m) == pytest.approx(0.0, abs=1e-28)
        assert M.sqrt_c_l2(k * I, I, m) == pytest.approx(0.0, abs=1e-20 * k)
        assert M.sqrt_c_l2(I, k * I, m) == pytest.approx(0.0, abs=1e-20)


def test_closed_forms(images):
    O, I, m = images
    assert M.l2(O, I, m) == pytest.approx(float(np.sum(m * (O - I) ** 2)), rel=1e-14)
    sO, sI = np.sqrt(O), np.sqrt(I)
    # sqrt_c_l2: c is the masked least-squares scale of sqrt I onto sqrt O, so D is the residual of
    # that projection: sum m (sqrt O)^2 - <sqrt I, sqrt O>^2 / ||sqrt I||^2 for a 0/1 mask
    proj = np.sum(m * sO * sO) - np.sum(m * sO * sI) ** 2 / np.sum(m * sI * sI)
    assert M.sqrt_c_l2(O, I, m) == pytest.approx(proj, rel=1e-12)
    # sqrt_unit_l2 of unit vectors: 2 - 2 cos(angle) for a 0/1 mask
    cos = np.sum(m * sO * sI) / np.sqrt(np.sum(m * sO * sO) * np.sum(m * sI * sI))
    assert M.sqrt_unit_l2(O, I, m) == pytest.approx(2 - 2 * cos, rel=1e-12)


def test_mask_weights_twice_in_the_distance():
    """A mask value w weights the norms once and the squared difference twice (w^2), as the reference
    multiplies the absolute difference by the mask before squaring."""
    O, I = np.array([4.0, 9.0]), np.array([1.0, 1.0])
    w = np.array([1.0, 0.5])
    assert M.l2(O, I, w) == pytest.approx(9.0 + 0.25 * 64.0)
    assert M.l2(O, I, np.ones(2)) == pytest.approx(9.0 + 64.0)


def test_cube_root_forms(images):
    """cbrt_unit_l2 and cbrt_c_l2: scale invariance (O = k I gives 0), the same closed forms over cube
    roots, and a negative value kept with its sign (cbrt, not a root of |x|)."""
    O, I, m = images
    for k in (1.0, 5.3, 1e-3):
        assert M.cbrt_unit_l2(k * I, I, m) == pytest.approx(0.0, abs=1e-26)
        assert M.cbrt_c_l2(k * I, I, m) == pytest.approx(0.0, abs=1e-20)
    cO, cI = np.cbrt(O), np.cbrt(I)
    proj = np.sum(m * cO * cO) - np.sum(m * cO * cI) ** 2 / np.sum(m * cI * cI)
    assert M.cbrt_c_l2(O, I, m) == pytest.approx(proj, rel=1e-12)
    cos = np.sum(m * cO * cI) / np.sqrt(np.sum(m * cO * cO) * np.sum(m * cI * cI))
    assert M.cbrt_unit_l2(O, I, m) == pytest.approx(2 - 2 * cos, rel=1e-12)
    # cube roots differ from square roots: the two unit forms disagree on generic data
    assert abs(M.cbrt_unit_l2(O, I, m) - M.sqrt_unit_l2(O, I, m)) > 1e-6
    # (-8, 1) against (8, 1): cube roots (-2, 1) and (2, 1), unit vectors apart by 4 * (2/sqrt 5)^2
    assert M.cbrt_unit_l2(np.array([-8.0, 1.0]), np.array([8.0, 1.0]), np.ones(2)) == pytest.approx(16 / 5)


def test_l2_keeps_non_finite_terms_and_the_objective_penalises_them():
    """AbsoluteDifferenceSquare (distance_functions.hpp:125-143) sums every term, so a NaN model pixel
    makes the distance NaN, and the fit objective then takes the 1e8 penalty
    (objective_func_hipgisaxs.cpp:297-312); the square-root forms drop the pixel instead."""
    O = np.array([1.0, 2.0, 3.0])
    I = np.array([1.5, np.nan, 2.0])
    m = np.ones(3)
    assert np.isnan(M.l2(O, I, m))
    assert np.isfinite(M.sqrt_c_l2(O, I, m))
    import math

    d = M.l2(O, I, m)
    assert (d if math.isfinite(d) else 1e8) == 1e8
