"""chi^2 forward/backward pair (hgx_chi2_fwd / hgx_chi2_bwd) for cylinders on an hcp lattice:
the forward chi^2 against the oracle composition (FF, SF, DWBA, chi^2 in FP64 on the CPU), the
adjoint gradient against central finite differences of the oracle chi^2 (h = 1e-6 relative) and
of the GPU forward."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 12.0)
AI = np.radians(0.15)
UA, UB, UC = np.array([1.0, 0, 0]), np.array([0.5, 0.866, 0]), np.array([0.5, 0.2887, 0.8165])
REP = (20, 20, 1)
DN2 = 2 * 9.4e-6 + 2j * 3.37e-8


def _rx(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


def _oracle_I(O, g, fc, theta, shape="cylinder", rot0=np.eye(3)):
    R, H, a, t, sig, sc = theta
    rot = _rx(t) @ rot0
    dims = [([R], [1.0]), ([H], [1.0])] if shape == "cylinder" else [([R], [1.0])]
    F = DN2 * O.ff_analytic(shape, dims, g, rot)
    S = O.sf_lattice(a * UA, a * UB, a * UC, REP, (0, 0, 0), g, dw_sigma=sig, rot=rot)
    return sc * O.dwba_intensity(g.npix, g.nq // g.npix, fc, S, F, 1.0)


def _chi2(I, obs, mask):
    return float(np.sum(mask * (I - obs) ** 2 / np.maximum(obs, 1.0)))


@pytest.fixture(scope="module")
def setup(oracle_mod):
    O = oracle_mod
    g = O.qgrid_detector(48, 40, 0.172 * 10, 2000.0, 23.5, 0.0, AI, K0)
    g = O.qz_extended(g, K0, AI, 0j)
    fc = O.fresnel(g, [3.39e-6], [7.37e-8], [0.0], 0, K0, AI)
    theta_true = np.array([4.0, 12.0, 15.0, 0.02, 0.5, 1.0])
    I_true = _oracle_I(O, g, fc, theta_true)
    rng = np.random.default_rng(20261203 + 4)
    obs = rng.poisson(1e4 * I_true / I_true.max()).astype(float) * (I_true.max() / 1e4)
    mask = (rng.uniform(size=g.npix) > 0.05).astype(float)
    return g, fc, obs, mask, theta_true


def _dev(x):
    import torch

    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = np.stack([x.real, x.imag], -1).ravel()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


@pytest.mark.parametrize("shape", ["cylinder", "sphere"])
def test_chi2_forward_and_adjoint(gpu_ctx, oracle_mod, setup, shape):
    from hipgisaxs_amd.api import FitPlan

    O = oracle_mod
    g, fc, obs, mask, theta_true = setup
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    fcd, obsd, maskd = _dev(fc), _dev(obs), _dev(mask)
    plan = FitPlan(gpu_ctx, shape, UA, UB, UC, REP, obsd, fc=fcd, mask=maskd, dn2=DN2)
    theta = theta_true * np.array([1.03, 0.97, 1.01, 1.0, 1.2, 0.9]) + np.array([0, 0, 0, 0.01, 0, 0])
    if shape == "sphere":
        theta[0] = 5.0
    f_or = lambda th: _chi2(_oracle_I(O, g, fc, th, shape), obs, mask)
    c_ref = f_or(theta)
    c_fwd = plan.fwd(theta)
    assert abs(c_fwd - c_ref) <= 1e-10 * abs(c_ref)
    c_bwd, grad = plan.bwd(theta)
    assert abs(c_bwd - c_ref) <= 1e-10 * abs(c_ref)
    # central differences of the oracle chi^2, h = 1e-6 relative
    fd = np.zeros(6)
    for j in range(6):
        if shape == "sphere" and j == 1:
            continue
        h = 1e-6 * max(abs(theta[j]), 1e-2)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (f_or(tp) - f_or(tm)) / (2 * h)
    print(shape, "grad", grad, "fd", fd)
    assert rel_l2(grad, fd) <= 1e-6
    for j in range(6):
        assert abs(grad[j] - fd[j]) <= 1e-5 * max(abs(fd[j]), 1e-8 * np.abs(fd).max()), (j, grad[j], fd[j])
    # and central differences of the GPU forward itself
    fdg = np.zeros(6)
    for j in (0, 2, 3, 4, 5):
        h = 1e-6 * max(abs(theta[j]), 1e-2)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fdg[j] = (plan.fwd(tp) - plan.fwd(tm)) / (2 * h)
    assert np.allclose(grad[[0, 2, 3, 4, 5]], fdg[[0, 2, 3, 4, 5]], rtol=1e-5, atol=1e-8 * np.abs(fdg).max())
    # the model image at the true parameters reproduces the oracle image
    import torch

    Im = torch.empty(g.npix, dtype=torch.float64, device="cuda")
    plan.fwd(theta_true if shape == "cylinder" else theta, I_model=Im)
    ref_I = _oracle_I(O, g, fc, theta_true if shape == "cylinder" else theta, shape)
    assert rel_l2(Im.cpu().numpy(), ref_I) <= 1e-10
    plan.release()


@pytest.mark.parametrize("metric", ["chi2", "l2", "sqrt_unit_l2", "sqrt_c_l2"])
def test_distance_metrics_and_their_adjoints(gpu_ctx, oracle_mod, setup, metric):
    """The reference's distance functors as objectives: value against the oracle's restatement on the
    oracle image, gradient against central differences of that oracle objective."""
    from hipgisaxs_amd.api import FitPlan
    from oracle import metrics as M

    O = oracle_mod
    g, fc, obs, mask, theta_true = setup
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    plan = FitPlan(gpu_ctx, "cylinder", UA, UB, UC, REP, _dev(obs), fc=_dev(fc), mask=_dev(mask), dn2=DN2)
    plan.metric(metric)
    theta = theta_true * np.array([1.02, 0.98, 1.005, 1.0, 1.1, 0.9]) + np.array([0, 0, 0, 0.01, 0, 0])
    f_or = lambda th: M.METRICS[metric](obs, _oracle_I(O, g, fc, th), mask)
    ref = f_or(theta)
    val = plan.fwd(theta)
    print(metric, "value", val, ref)
    assert abs(val - ref) <= 1e-10 * abs(ref)
    v2, grad = plan.bwd(theta)
    assert abs(v2 - ref) <= 1e-10 * abs(ref)
    fd = np.zeros(6)
    for j in range(6):
        h = 1e-6 * max(abs(theta[j]), 1e-2)
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd[j] = (f_or(tp) - f_or(tm)) / (2 * h)
    print(metric, "grad", grad, "fd", fd)
    scale = np.abs(fd).max()
    for j in range(6):
        assert abs(grad[j] - fd[j]) <= 1e-5 * max(abs(fd[j]), 1e-6 * scale), (metric, j, grad[j], fd[j])
    if metric in ("sqrt_unit_l2", "sqrt_c_l2"):  # scale-invariant objectives: no gradient along the scale
        assert abs(grad[5]) <= 1e-9 * scale
    plan.release()


def test_population_evaluation_matches_single_calls(gpu_ctx, setup):
    from hipgisaxs_amd.api import FitPlan

    g, fc, obs, mask, theta_true = setup
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    plan = FitPlan(gpu_ctx, "cylinder", UA, UB, UC, REP, _dev(obs), fc=_dev(fc), mask=_dev(mask), dn2=DN2)
    rng = np.random.default_rng(3)
    pop = theta_true * (1 + 0.05 * rng.standard_normal((9, 6)))
    pop[:, 3] = 0.02 + 0.01 * rng.standard_normal(9)
    for metric in ("chi2", "sqrt_c_l2"):
        plan.metric(metric)
        many = plan.fwd_many(pop)
        single = np.array([plan.fwd(t) for t in pop])
        assert np.array_equal(many, single), (metric, many, single)
    plan.release()


@pytest.mark.parametrize("metric,free", [("chi2", (1, 1, 1, 1, 1, 1)), ("sqrt_c_l2", (1, 1, 1, 1, 1, 0))])
def test_lbfgs_fit_recovers_the_parameters(gpu_ctx, setup, metric, free):
    """Bounded L-BFGS with the adjoint gradient, started inside the basin of the true parameters (the
    lattice peaks make the objective multimodal: a start 3 % off in the lattice constant already ends in
    another minimum, as any local method would): the objective ends at or below its value at the true
    parameters (the noisy data's optimum lies near them) and the parameters come back to within the
    noise."""
    from hipgisaxs_amd.api import FitPlan

    g, fc, obs, mask, theta_true = setup
    gpu_ctx.qgrid_set(g.ncol, g.nrow, g.qx, g.qy, g.qz, g.alpha_f)
    plan = FitPlan(gpu_ctx, "cylinder", UA, UB, UC, REP, _dev(obs), fc=_dev(fc), mask=_dev(mask), dn2=DN2)
    plan.metric(metric)
    start = theta_true * np.array([1.01, 0.99, 1.003, 1.0, 1.05, 1.0 if not free[5] else 0.97])
    start[3] = 0.015
    lo = np.array([2.0, 6.0, 10.0, -0.1, 0.0, 0.1])
    hi = np.array([6.0, 20.0, 20.0, 0.1, 2.0, 10.0])
    f_true = plan.fwd(theta_true)
    f0 = plan.fwd(start)
    th, f, it = plan.lbfgs(start, free=free, lo=lo, hi=hi, iterations=200, tol=1e-10)
    print(metric, "f0", f0, "f", f, "f_true", f_true, "iterations", it, "theta", th)
    assert f <= f_true * (1 + 1e-6)
    assert f < 0.2 * f0
    rel = np.abs(th - theta_true) / np.abs(theta_true)
    assert rel[0] < 0.01 and rel[1] < 0.02 and rel[2] < 0.01, rel
    assert np.all(th >= lo) and np.all(th <= hi)
    # the result is a stationary point of the free parameters
    _, grad = plan.bwd(th)
    assert np.all(np.abs(grad * (hi - lo) * np.array(free)) <= 1e-4 * abs(f) + 1e-12)
    plan.release()
