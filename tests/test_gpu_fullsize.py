"""Parity at the workloads' full sizes, where the CPU oracle would take minutes: FP32C against the plain
FP64 GPU path (itself matched to the oracle at 1e-10 in the other tests) on a 1024^2 grid, and
size-independent identities (Friedel symmetry of real-q form factors, the translation phase)."""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

K0 = 2 * math.pi / (1.23984 / 10.0)
AI = math.radians(0.2)


def test_numeric_fp32c_full_size_grid():
    """A 50 000-triangle polyhedron on the 1024^2 Born grid (the largest shape of the batched workload):
    matrix-core FP32C (FP32 and split-fp16 operands) and per-pair FP32C against plain FP64, and
    Friedel symmetry F(-q) = conj F(q)."""
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import PRODUCT_DEFAULT, Context, tune_product

    rng = np.random.default_rng(20261203)
    tri = meshes.convex_polyhedron(rng, np.array([15.0, 11.0, 7.0]), 50000)
    c = Context(0)
    try:
        c.qgrid_explicit(1024, 1024, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
        mid = c.upload_mesh(tri, "exact9")
        f64 = c.as_complex(c.ff_numeric(mid, "approx", "fp64")).cpu().numpy()
        fva = c.as_complex(c.ff_numeric(mid, "approx", "fp32c", flags=4)).cpu().numpy()
        try:
            tune_product(c, 16, 3, "fp32")
            f32 = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy()
            f16 = {}
            for variant in ((16, 4, "f16x2", 1, True, False), (64, 4, "f16x2", 2, True, False),
                            (64, 4, "f16x2", 4, True, False), (64, 4, "f16x2", 4, False, False),
                            (64, 4, "f16x2", 2, True, True), (64, 4, "f16x2", 4, False, True),
                            (64, 4, "f16x2", 4, True, True)):
                tune_product(c, *variant)
                f16[variant] = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy()
        finally:
            tune_product(c, *PRODUCT_DEFAULT)
        e_mfma, e_valu = rel_l2(f32, f64), rel_l2(fva, f64)
        e_16 = {"x".join(map(str, k)): rel_l2(v, f64) for k, v in f16.items()}
        print(f"1024^2 x 50k: mfma {e_mfma:.2e} valu {e_valu:.2e} split-fp16 mfma",
              {k: f"{v:.2e}" for k, v in e_16.items()})
        assert e_mfma <= 1e-5 and e_valu <= 1e-5 and max(e_16.values()) <= 1e-5
        # Friedel: a grid symmetric in qy with qz -> -qz gives the conjugate
        c.qgrid_explicit(512, 256, -2.0, -2.0, 2.0, 0.0, 0.0, 1e3)  # qz in [-2, 0]
        lo = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy().reshape(256, 512)
        c.qgrid_explicit(512, 256, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)  # qz in [0, 2]
        hi = c.as_complex(c.ff_numeric(mid, "approx", "fp32c")).cpu().numpy().reshape(256, 512)
        # row i of hi has qz = 2 (255 - i)/255; row i of lo has qz = -2 i/255; columns qy -> -qy reversed
        mirrored = np.conj(lo[::-1, ::-1])
        e = rel_l2(hi, mirrored)
        print(f"Friedel relL2 {e:.2e}")
        assert e <= 1e-5
    finally:
        c.close()


def test_numeric_translation_phase_full_size():
    """F of a translated mesh = exp(i q.t) F, on the 1024^2 DWBA grid (complex branches)."""
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Context

    rng = np.random.default_rng(4)
    tri = meshes.convex_polyhedron(rng, np.array([8.0, 6.0, 5.0]), 8000)
    t = np.array([3.0, -7.5, 2.25])
    moved = tri + np.tile(t, 3)
    c = Context(0)
    try:
        c.qgrid_explicit(1024, 1024, -2.0, 0.0, 2.0, 2.0, AI, K0)
        c.qz_extended(K0, AI, 3e-6 + 2e-8j)
        m0 = c.upload_mesh(tri, "exact9")
        m1 = c.upload_mesh(moved, "exact9")
        a = c.as_complex(c.ff_numeric(m0, "approx", "fp32c", transvec=t)).cpu().numpy()
        b = c.as_complex(c.ff_numeric(m1, "approx", "fp32c")).cpu().numpy()
        e = rel_l2(b, a)
        print(f"translation relL2 {e:.2e}")
        assert e <= 1e-5
    finally:
        c.close()


def test_lattice_fp32c_full_size():
    """The sphere-lattice workload at full size (1024^2 DWBA, 21 nodes, 8 azimuths): FP32C vs FP64."""
    import torch

    from hipgisaxs_amd.api import Context, lattice_intensity, param_distribution

    k0 = 2 * math.pi / (1.23984 / 12.0)
    ai = math.radians(0.15)
    c = Context(0)
    try:
        c.qgrid_detector(1024, 1024, 0.172, 2000.0, 511.5, 0.0, ai, k0)
        c.qz_extended(k0, ai, 0j)
        fc = c.fresnel([3.39e-6], [3.55e-8], [0.0], 0, k0, ai)
        v, w = param_distribution(3.0, 7.0, 21, 2, 5.0, 0.5)
        a = 20.0
        rots = np.stack([np.array([[math.cos(p), -math.sin(p), 0], [math.sin(p), math.cos(p), 0], [0, 0, 1]])
                         for p in (j * math.pi / 4 for j in range(8))])
        out = {}
        for prec in ("fp32c", "fp64"):
            I = torch.empty(1024 * 1024, dtype=torch.float64, device="cuda")
            lattice_intensity(c, "sphere", [(v, w / w.sum())], (a, 0, 0), (0, a, 0), (a / 2, a / 2, a / 2),
                              (10, 10, 5), rots, None, dw_sigma=0.5, dn2=-(2 * 6.53e-6 + 2j * 1.6e-8),
                              precision=prec, fc=fc, out=I)
            out[prec] = I.cpu().numpy()
        e = rel_l2(out["fp32c"], out["fp64"])
        print(f"lattice 1024^2 relL2 {e:.2e}")
        assert e <= 1e-5
    finally:
        c.close()
