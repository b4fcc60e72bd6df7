"""Batched trapezoid-grating fit (hgx_grating_*): device mesh, FP32C forward on the matrix cores, chi^2
and its adjoint gradient, on-device L-BFGS.

Parity: model images and chi^2 against a host reference built from the oracle's FP64 form factor of
the same mesh (meshes.trapezoid_line); gradients against the oracle's record adjoint (pinned by finite
differences in test_oracle_adjoint.py) chained through finite-difference mesh derivatives, and against
central differences of the plan's own chi^2."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi / (1.23984 / 10.0)
AI = np.radians(0.2)
MODEL = dict(length=40.0, pitch=100.0, periods=5, n_bottom=16, n_side=8, n_top=16, n_length=8)
DN2 = 2e-5 + 1e-7j


def _dirichlet(x, n):
    x = x - 2 * np.pi * np.rint(x / (2 * np.pi))
    s = np.sin(x / 2)
    out = np.full_like(x, float(n))
    nz = np.abs(s) >= 1e-12
    out[nz] = np.sin(n * x[nz] / 2) / s[nz]
    return out


class Host:
    """FP64 host reference of the grating model on an oracle grid."""

    def __init__(self, O, g, fc, obs, model=MODEL, dn2=DN2):
        self.O, self.g, self.fc, self.obs, self.m, self.dn2 = O, g, fc, obs, model, dn2
        self.nb = g.nq // g.npix

    def mesh(self, th):
        from hipgisaxs_amd import meshes

        m = self.m
        return meshes.trapezoid_line(th[0], th[1], th[2], m["length"], m["n_bottom"], m["n_side"], m["n_top"],
                                     m["n_length"])

    def parts(self, th):
        O, g = self.O, self.g
        tri = self.mesh(th)
        rec = O.tri9_to_rec8(tri)
        F = O.ff_approx(rec, g).reshape(self.nb, g.npix)
        fcb = self.fc.reshape(self.nb, g.npix) if self.fc is not None else np.ones_like(F)
        A = self.dn2 * (fcb * F).sum(0)
        sf = _dirichlet(g.qy * self.m["pitch"], self.m["periods"])
        I = th[3] * sf ** 2 * np.abs(A) ** 2
        wb = th[0] + 2 * th[2] / np.tan(th[1])
        R = np.sqrt(self.m["length"] ** 2 / 4 + max(wb, th[0]) ** 2 / 4 + th[2] ** 2 / 4)
        qn = np.sqrt(np.abs(g.q_rotated()) ** 2).reshape(3, self.nb, g.npix)
        qn = np.sqrt((qn ** 2).sum(0))
        near0 = (qn * R < 0.1).any(0)
        wgt = np.where(near0, 0.0, 1.0) / np.maximum(self.obs, 1.0)
        return tri, rec, A, fcb, sf, I, wgt

    def chi2(self, th):
        _, _, _, _, _, I, wgt = self.parts(th)
        return float(np.sum(wgt * (I - self.obs) ** 2))

    def grad(self, th, h=1e-6):
        O, g = self.O, self.g
        tri, rec, A, fcb, sf, I, wgt = self.parts(th)
        res = I - self.obs
        W = (4 * wgt * res * th[3] * sf ** 2 * np.conj(A) * self.dn2)[None, :] * fcb
        ga = O.ff_approx_adjoint(rec, g, W.ravel())
        A0 = rec[:, :1] * rec[:, 1:4]
        out = np.zeros(4)
        for j in range(3):
            tp, tm = np.array(th, float), np.array(th, float)
            dh = h * max(1.0, abs(th[j]))
            tp[j] += dh
            tm[j] -= dh
            rp, rm = O.tri9_to_rec8(self.mesh(tp)), O.tri9_to_rec8(self.mesh(tm))
            dA = (rp[:, :1] * rp[:, 1:4] - rm[:, :1] * rm[:, 1:4]) / (2 * dh)
            dc = (rp[:, 4:7] - rm[:, 4:7]) / (2 * dh)
            out[j] = np.sum(ga[:, :3] * dA) + np.sum(ga[:, 3:] * dc)
        out[3] = np.sum(2 * wgt * res * I / th[3])
        assert A0.shape[0] == len(tri)
        return out


def _setup(O, ncol, nrow, dnl):
    import torch

    from hipgisaxs_amd.api import Context

    c = Context(0)
    c.qgrid_explicit(ncol, nrow, -1.0, 0.0, 1.0, 1.5, AI, K0)
    c.qz_extended(K0, AI, dnl)
    g = O.qz_extended(O.qgrid_explicit(ncol, nrow, -1.0, 0.0, 1.0, 1.5, AI, K0), K0, AI, dnl)
    fc = O.fresnel(g, [4.88e-6], [7.37e-8], [0.0], 0, K0, AI)
    fc_d = torch.from_numpy(np.stack([fc.real, fc.imag], -1).ravel()).cuda()
    return c, g, fc, fc_d


THETA_TRUE = np.array([20.0, np.radians(82.0), 28.0, 1.0])


@pytest.mark.parametrize("dnl", [0j, 3e-6 + 2e-8j])
def test_grating_forward_chi2_and_gradient(oracle_mod, dnl):
    import torch

    from hipgisaxs_amd.api import GratingPlan

    O = oracle_mod
    c, g, fc, fc_d = _setup(O, 48, 40, dnl)
    try:
        assert c.is_product()
        # observed image: the host model at THETA_TRUE, scaled to a peak of 1e4 counts
        host0 = Host(O, g, fc, np.zeros(g.npix))
        I1 = host0.parts(THETA_TRUE)[5]
        th_true = THETA_TRUE.copy()
        th_true[3] = 1e4 / I1.max()
        obs = th_true[3] * I1
        host = Host(O, g, fc, obs)
        obs_d = torch.from_numpy(obs).cuda()
        thetas = np.array([th_true * [1.05, 1.0, 0.95, 1.1],
                           th_true * [0.9, 1.02, 1.08, 0.8],
                           th_true * [1.0, 0.97, 1.0, 1.0]])
        plan = GratingPlan(c, 3, obs_d, fc=fc_d, dn2=DN2, **MODEL)
        plan.set_theta(thetas)
        plan.eval(use_graph=False)
        th, chi2, grad = plan.get()
        for k in range(3):
            I_gpu = plan.intensity(k).cpu().numpy()
            _, _, _, _, _, I_ref, wgt = host.parts(thetas[k])
            keep = wgt > 0
            e_i = rel_l2(I_gpu[keep], I_ref[keep])
            c_ref = host.chi2(thetas[k])
            g_ref = host.grad(thetas[k])
            e_c = abs(chi2[k] - c_ref) / c_ref
            e_g = np.max(np.abs(grad[k] - g_ref) / np.maximum(np.abs(g_ref), 1e-3 * np.abs(g_ref).max()))
            print(f"dnl={dnl} cand {k}: I relL2 {e_i:.2e}, chi2 {e_c:.2e}, grad {e_g:.2e}; "
                  f"grad {grad[k]} ref {g_ref}")
            assert e_i <= 1e-5
            assert e_c <= 1e-5
            assert e_g <= 1e-4
        # graph replay gives the same numbers
        plan.eval(use_graph=True)
        plan.eval(use_graph=True)
        _, chi2_g, grad_g = plan.get()
        assert np.allclose(chi2_g, chi2, rtol=1e-12) and np.allclose(grad_g, grad, rtol=1e-12)
        plan.release()
    finally:
        c.close()


def test_grating_gradient_matches_central_differences(oracle_mod):
    import torch

    from hipgisaxs_amd.api import GratingPlan

    O = oracle_mod
    c, g, fc, fc_d = _setup(O, 40, 32, 0j)
    try:
        host0 = Host(O, g, fc, np.zeros(g.npix))
        I1 = host0.parts(THETA_TRUE)[5]
        scale = 1e4 / I1.max()
        obs_d = torch.from_numpy(scale * I1).cuda()
        th0 = THETA_TRUE * [1.04, 0.99, 0.96, scale * 1.1]
        rel = np.array([1e-3, 1e-3, 1e-3, 1e-3])
        thetas = [th0]
        for j in range(4):
            for sgn in (+1, -1):
                t = th0.copy()
                t[j] *= 1 + sgn * rel[j]
                thetas.append(t)
        plan = GratingPlan(c, len(thetas), obs_d, fc=fc_d, dn2=DN2, **MODEL)
        plan.set_theta(np.array(thetas))
        plan.eval()
        _, chi2, grad = plan.get()
        for j in range(4):
            fd = (chi2[1 + 2 * j] - chi2[2 + 2 * j]) / (2 * rel[j] * th0[j])
            print(f"param {j}: adjoint {grad[0][j]:.6e} central difference {fd:.6e}")
            assert abs(fd - grad[0][j]) <= 2e-3 * abs(grad[0][j])
        plan.release()
    finally:
        c.close()


def test_grating_lbfgs_recovers_parameters(oracle_mod):
    import torch

    from hipgisaxs_amd.api import GratingPlan

    O = oracle_mod
    c, g, fc, fc_d = _setup(O, 40, 32, 0j)
    try:
        host0 = Host(O, g, fc, np.zeros(g.npix))
        I1 = host0.parts(THETA_TRUE)[5]
        scale = 1e4 / I1.max()
        truth = THETA_TRUE.copy()
        truth[3] = scale
        obs_d = torch.from_numpy(scale * I1).cuda()
        starts = np.array([truth * [1.06, 0.985, 0.95, 1.2], truth * [0.95, 1.0, 1.05, 0.85]])
        plan = GratingPlan(c, 2, obs_d, fc=fc_d, dn2=DN2, **MODEL)
        plan.set_theta(starts)
        plan.eval()
        _, chi0, _ = plan.get()
        plan.lbfgs(60)
        th, chi2, _ = plan.get()
        print("chi2", chi0, "->", chi2)
        print("theta", th, "truth", truth)
        for k in range(2):
            assert chi2[k] <= 1e-3 * chi0[k]
            assert np.all(np.abs(th[k] - truth) <= 0.01 * np.abs(truth))
        plan.release()
    finally:
        c.close()


def test_grating_full_size_image_matches_fp64_path():
    """The grating workload at full size (106 912 triangles, 1024^2 grid, four DWBA branches): the plan's
    FP32C model image against the plain FP64 GPU path on the same mesh built on the host."""
    import math

    import torch

    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Context, GratingPlan

    k0 = 2 * math.pi / (1.23984 / 10.0)
    ai = math.radians(0.2)
    model = dict(length=100.0, pitch=100.0, periods=200, n_bottom=64, n_side=40, n_top=64, n_length=256)
    dn2 = 2e-5 + 1e-7j
    th = np.array([20.0, math.radians(82.0), 40.0, 3.0])
    c = Context(0)
    try:
        c.qgrid_explicit(1024, 1024, -2.0, 0.0, 2.0, 2.0, ai, k0)
        c.qz_extended(k0, ai, 0j)
        fc = c.fresnel([4.88e-6], [7.37e-8], [0.0], 0, k0, ai)
        zero = torch.zeros(1024 * 1024, dtype=torch.float64, device="cuda")
        plan = GratingPlan(c, 1, zero, fc=fc, dn2=dn2, **model)
        plan.set_theta(th[None, :])
        I32 = plan.intensity(0).cpu().numpy()
        plan.release()
        tri = meshes.trapezoid_line(th[0], th[1], th[2], model["length"], model["n_bottom"], model["n_side"],
                                    model["n_top"], model["n_length"])
        mid = c.upload_mesh(tri, "exact9")
        F = c.as_complex(c.ff_numeric(mid, "approx", "fp64")).reshape(4, -1)
        A = (dn2 * (c.as_complex(fc).reshape(4, -1) * F).sum(0)).cpu().numpy()
        qx, qy, qz, _ = c.qgrid_get()
        sf = _dirichlet(qy * model["pitch"], model["periods"])
        I64 = th[3] * sf ** 2 * np.abs(A) ** 2
        wb = th[0] + 2 * th[2] / np.tan(th[1])
        R = np.sqrt(model["length"] ** 2 / 4 + wb ** 2 / 4 + th[2] ** 2 / 4)
        qn = np.sqrt(qx[None] ** 2 + qy[None] ** 2 + np.abs(qz.reshape(4, -1)) ** 2)
        keep = ~(qn * R < 0.1).any(0)
        e = rel_l2(I32[keep], I64[keep])
        print(f"grating 1024^2 x {len(tri)}: relL2 {e:.2e}")
        assert e <= 1e-5
    finally:
        c.close()


@pytest.mark.parametrize("metric", ["l2", "sqrt_unit_l2", "sqrt_c_l2"])
def test_grating_distance_metrics(oracle_mod, metric):
    """The reference's distance functors as the batched fit's objective: values against the oracle's
    functors on the host FP64 model image (with the plan's small-|q| exclusion as the mask), gradients
    against central differences (h = 1e-5 relative) of that FP64 host objective, and graph replay."""
    import torch

    from hipgisaxs_amd.api import GratingPlan
    from oracle import metrics as M

    O = oracle_mod
    c, g, fc, fc_d = _setup(O, 40, 32, 0j)
    try:
        host0 = Host(O, g, fc, np.zeros(g.npix))
        I1 = host0.parts(THETA_TRUE)[5]
        scale = 1e4 / I1.max()
        obs = scale * I1 * (1 + 0.05 * np.sin(np.arange(g.npix)))  # data not exactly of the model's form
        host = Host(O, g, fc, obs)
        th0 = THETA_TRUE * [1.04, 0.99, 0.96, scale * 1.1]

        def f_host(th):
            parts = host.parts(th)
            return M.METRICS[metric](obs, parts[5], (parts[6] > 0).astype(float))

        plan = GratingPlan(c, 1, torch.from_numpy(obs).cuda(), fc=fc_d, dn2=DN2, **MODEL)
        plan.metric(metric)
        plan.set_theta(th0[None])
        plan.eval()
        _, val, grad = plan.get()
        ref = f_host(th0)
        print(metric, "value", val[0], "host", ref)
        assert abs(val[0] - ref) <= 1e-4 * abs(ref)
        fd = np.zeros(4)
        for j in range(4):
            h = 1e-5 * th0[j]
            tp, tm = th0.copy(), th0.copy()
            tp[j] += h
            tm[j] -= h
            fd[j] = (f_host(tp) - f_host(tm)) / (2 * h)
        print(metric, "adjoint", grad[0], "host central differences", fd)
        big = np.abs(fd * th0).max()
        for j in range(4):
            assert abs(fd[j] - grad[0][j]) * th0[j] <= 2e-3 * max(abs(fd[j]) * th0[j], 1e-3 * big), j
        if metric != "l2":  # scale-invariant objectives
            assert abs(grad[0][3]) * th0[3] <= 1e-6 * big
        plan.eval(use_graph=True)
        plan.eval(use_graph=True)
        _, val_g, grad_g = plan.get()
        assert np.allclose(val_g, val, rtol=1e-12) and np.allclose(grad_g, grad, rtol=1e-12)
        plan.release()
    finally:
        c.close()


def test_grating_plan_refuses_a_replaced_grid(oracle_mod):
    """A plan keeps the device addresses of its grid: after the grid is replaced (or extended) eval,
    lbfgs and intensity refuse before any launch, with a message that says what to do."""
    import torch

    from hipgisaxs_amd.api import GratingPlan, HgxError

    O = oracle_mod
    c, g, fc, fc_d = _setup(O, 32, 24, 0j)
    try:
        obs_d = torch.ones(g.npix, dtype=torch.float64, device="cuda")
        plan = GratingPlan(c, 1, obs_d, fc=fc_d, dn2=DN2, **MODEL)
        plan.set_theta(THETA_TRUE[None, :])
        plan.eval(use_graph=False)
        torch.cuda.synchronize()
        c.qgrid_explicit(32, 24, -1.0, 0.0, 1.0, 1.5, AI, K0)  # same size, new arrays
        for call in (lambda: plan.eval(use_graph=False), lambda: plan.eval(use_graph=True),
                     lambda: plan.lbfgs(1), lambda: plan.intensity(0)):
            with pytest.raises(HgxError, match="create the plan again"):
                call()
        plan.release()
    finally:
        c.close()
