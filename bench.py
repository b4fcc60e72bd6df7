#!/usr/bin/env python3
"""Benchmark of the numeric form-factor hot path on one MI355X.

Workload (default): 64 random convex polyhedra (seeded; points on ellipsoids with semi-axes
U[3,15] nm, convex hull), triangle counts log-uniform in [500, 50000], each with its own
Born q-grid qy in [-2, 2], qz in [0, 2] nm^-1 of side 128/256/512/1024. One step = the
approximate-triangle form factor of all 64 shapes on their grids. The headline value is the
FP32-compensated path's q-point x triangle pairs per second; the plain FP64 GPU path is timed
beside it on the same inputs, with relL2 of FP32C against FP64 on the full outputs and against
the CPU oracle on a sample.

Prints ONE JSON line (rank 0). Only one GPU is used: --gpus N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "q-pt·triangles/s + χ²-grad evals/s, FP32-comp vs plain FP64 GPU @ relL2≤1e-5"
SEED = 20261203
PEAK_FP32 = 157.3e12  # MI355X dense FP32 peak (matrix cores; also packed vector FMA)
PEAK_FP64 = 78.6e12
FLOP_PER_PAIR = 33  # approximate-triangle pair, direct form: q.n 10 + q.c 10 + phase 2 + scale 3 + cmul 6 + acc 2
# Product-grid form: two complex multiply-adds per (q, triangle) pair, each done as 3 (Gauss) or 4
# real multiply-adds of 2 flop on the matrix cores: 12 or 16 flop per pair.
def flop_per_pair_product(mults, operands="fp32"):
    # split fp16: each real product is three fp16 products (high x high, low x high, high x low)
    return 2 * mults * 2 * (3 if operands == "f16x2" else 1)


PEAK_FP16 = 2.5e15  # MI355X dense fp16 matrix-core peak


# HBM bytes per launch of the product kernel on this workload and of the finishing kernel that
# follows it, per work-unit shape, from rocprofv3 --pmc FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000
# (scales from a 256 MiB device copy) in KiB per dispatch (profiles/k_ff_product16s_64polyhedra_batch.md).
PRODUCT_TRAFFIC_BYTES = {(64, 4, "f16x2", 2, True, True): ((85599.1 * 2 + 413824.0) * 1024,
                                                           (522742.7 * 2 + 306688.0) * 1024)}


def ALGO_BYTES(shapes):
    """Algorithmic HBM bytes of one step of the product kernel: every triangle record once (32 B, FP32,
    padded to 128-triangle blocks) and every complex FP64 output once (16 B per q-point)."""
    return float(sum(32 * 128 * math.ceil(s["ntri"] / 128) + 16 * s["side"] ** 2 for s in shapes))


def intensity(z):
    """|F|^2 of a complex image stored as interleaved float64."""
    v = z.view(-1, 2)
    return v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]


class pixel_errors:
    """Per-pixel accuracy of FP32C intensity images I against FP64 ones I64, over many images:
    r = |I - I64| / max(I64, 1e-10 max I64) (the floor keeps exact nodes of F, where I64 -> 0, from
    dividing by rounding; ten decades below the peak), its maximum, median and the fractions of pixels
    above 1e-6, 1e-5 and 1e-4; and relL2 of log10 of the floored intensities (the error a fit on log
    intensities sees)."""

    FLOOR = 1e-10

    def __init__(self):
        self.max = 0.0
        self.n = 0
        self.over = {"1e-6": 0, "1e-5": 0, "1e-4": 0}
        self.med = []
        self.lnum = self.lden = 0.0

    def add(self, I, I64):
        import torch

        fl = self.FLOOR * float(I64.max())
        den = torch.clamp(I64, min=fl)
        r = (I - I64).abs() / den
        self.max = max(self.max, float(r.max()))
        self.n += r.numel()
        for k in self.over:
            self.over[k] += int((r > float(k)).sum())
        self.med.append(float(r.median()))
        l32 = torch.log10(torch.clamp(I, min=fl))
        l64 = torch.log10(den)
        self.lnum += float(torch.linalg.vector_norm(l32 - l64)) ** 2
        self.lden += float(torch.linalg.vector_norm(l64)) ** 2

    def result(self):
        return {"max_rel": self.max, "median_rel_of_images": float(np.median(self.med)) if self.med else None,
                "frac_rel_over": {k: v / max(self.n, 1) for k, v in self.over.items()},
                "relL2_log10I": math.sqrt(self.lnum / max(self.lden, 1e-300)), "floor_rel_to_peak": self.FLOOR,
                "pixels": self.n}


def make_workload(nshapes=64, seed=SEED + 3):
    from hipgisaxs_amd import meshes

    return meshes.c3_shapes(nshapes, seed)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(shapes, run_s=5.0, repeats=3):
    """Times the FP64 OpenMP oracle on a bounded sample of the workload (about 20 s of CPU work in all):
    the median-size shape among those on the largest grid, a band of its q-grid rows sized to about run_s
    per run, median of `repeats` runs; then the same loop as the reference writes it (a heap-allocated
    rotation per pair) on a band sized to about run_s, once."""
    import oracle as O

    big = max(d["side"] for d in shapes)
    pool = sorted((d for d in shapes if d["side"] == big), key=lambda d: d["tri"].shape[0])
    s = pool[len(pool) // 2]
    rec = O.tri9_to_rec8(s["tri"])
    side = s["side"]
    g = O.qgrid_explicit(side, side, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)

    def band(rows):
        return O.Grid(side, rows, g.qx[: rows * side], g.qy[: rows * side], g.qz[: rows * side] + 0j)

    def calibrate(fn):
        rows = 2
        while True:
            t0 = time.perf_counter()
            fn(rec, band(rows))
            dt = time.perf_counter() - t0
            if dt > 0.3 or rows >= side:
                return int(min(side, max(1, rows * run_s / max(dt, 1e-3))))
            rows = min(side, rows * 4)

    rows = calibrate(O.ff_approx)
    sub = band(rows)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = O.ff_approx(rec, sub)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    pairs = rows * side * rec.shape[0]
    rows_w = calibrate(O.ff_approx_as_written)
    t0 = time.perf_counter()
    O.ff_approx_as_written(rec, band(rows_w))
    dt_w = time.perf_counter() - t0
    pairs_w = rows_w * side * rec.shape[0]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": pairs / dt,
        "unit": "q-pt·triangles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 OpenMP oracle (rotation hoisted per q), approximate method, shape with {rec.shape[0]} "
                  f"triangles x {rows} rows of its {side}x{side} grid ({pairs:.3e} pairs), median of {repeats} "
                  f"runs of {dt:.1f} s",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "reference_as_written": {
            "value": pairs_w / dt_w,
            "sample": f"same shape, {rows_w} rows ({pairs_w:.3e} pairs, {dt_w:.1f} s), one heap-allocated "
                      f"rotation per pair as src/ff/cpu/ff_tri_cpu.cpp:233",
        },
    }, (s, rows, out)


def _det_grid(ctx, side, pixel_mm, alpha_i, k0, dwba=True):
    ctx.qgrid_detector(side, side, pixel_mm, 2000.0, (side - 1) / 2.0, 0.0, alpha_i, k0)
    if dwba:
        ctx.qz_extended(k0, alpha_i, 0j)


def exact32_flop_per_pair(g, tri, qpt=4):
    """Flop per pair of the FP32C exact kernel's series on grid g: 39 + 5 (NT - 4) for NT terms
    (phase, q.sn, q.d1, q.d2: 5 each; p, r: 6; series start: 3; complex combine: 10; each further
    pair of terms: 10; sin/cos not counted), NT chosen per wave (64 lanes x qpt q-points strided by
    256) from a^2 = max|q|^2 max|d|^2, here with the mesh-wide max|d| (an upper bound per block)."""
    npix, nq = len(g.qx), len(g.qz)
    q2 = np.tile(g.qx, nq // npix) ** 2 + np.tile(g.qy, nq // npix) ** 2 + g.qz.real ** 2
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3)
    bd2 = float((np.linalg.norm(t - t.mean(1)[:, None], axis=2).max()) ** 2)
    per_wg = 256 * qpt
    mix = {}
    for base in range(0, nq, per_wg):
        for w in range(4):
            idx = np.array([base + w * 64 + l + k * 256 for k in range(qpt) for l in range(64)])
            idx = np.minimum(idx, nq - 1)
            a2 = float(q2[idx].max()) * bd2
            nt = next((n for lim, n in ((0.25, 8), (1, 10), (4, 12), (9, 16), (16, 18)) if a2 <= lim), 0)
            mix[nt] = mix.get(nt, 0) + 1
    if 0 in mix:
        return None, mix  # divided-difference waves: no single flop model
    tot = sum(mix.values())
    return sum((39 + 5 * (n - 4)) * c for n, c in mix.items()) / tot, {str(k): v for k, v in mix.items()}


def bench_pyramid_fp64(reps=5):
    """The plain-method configuration at full size: the 2048-triangle square pyramid (base 20 nm, height
    10 nm) on the 256^2 q-space grid qy [-2, 2], qz [0, 2] with its four DWBA branches (10 keV,
    0.2 deg, Si substrate), FP64 approximate-triangle form factor and DWBA intensity, on the GPU and with
    the CPU oracle on the GPU box's host cores, and the relL2 between them (bar 1e-10)."""
    import torch
    import oracle as O
    from hipgisaxs_amd import meshes
    from hipgisaxs_amd.api import Context

    k0 = 2 * math.pi / (1.23984 / 10.0)
    ai = math.radians(0.2)
    tri = meshes.pyramid()
    g = O.qz_extended(O.qgrid_explicit(256, 256, -2.0, 0.0, 2.0, 2.0, ai, k0), k0, ai)
    fc = O.fresnel(g, [4.88e-6], [7.37e-8], [0.0], 0, k0, ai)
    ctx = Context(0)
    ctx.qgrid_explicit(256, 256, -2.0, 0.0, 2.0, 2.0, ai, k0)
    ctx.qz_extended(k0, ai, 0j)
    mid = ctx.upload_mesh(tri, "exact9")
    fcd = torch.from_numpy(np.stack([fc.real, fc.imag], -1).ravel()).cuda()
    ff = ctx.ff_numeric(mid, "approx", "fp64")
    I = ctx.dwba_intensity(ff, fc=fcd)
    torch.cuda.synchronize()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.ff_numeric(mid, "approx", "fp64", out=ff)
        ctx.dwba_intensity(ff, fc=fcd, out=I)
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t0) / reps
    k_approx = ctx.profile_end()
    # the exact (Green's theorem) triangle form factor of the same mesh and grid, FP64
    ffe = ctx.ff_numeric(mid, "exact", "fp64")
    torch.cuda.synchronize()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.ff_numeric(mid, "exact", "fp64", out=ffe)
    torch.cuda.synchronize()
    t_exact = (time.perf_counter() - t0) / reps
    k_exact = ctx.profile_end()
    # the FP32C exact method (FP32 series / divided differences per triangle) on the same inputs
    ffe32 = ctx.ff_numeric(mid, "exact", "fp32c")
    torch.cuda.synchronize()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.ff_numeric(mid, "exact", "fp32c", out=ffe32)
    torch.cuda.synchronize()
    t_exact32 = (time.perf_counter() - t0) / reps
    k_exact32 = ctx.profile_end()
    rel_e32 = float(torch.linalg.vector_norm(ffe32 - ffe) / torch.linalg.vector_norm(ffe))
    # the FP32C path on the same inputs: its intensity against the FP64 one, per pixel
    ff32 = ctx.ff_numeric(mid, "approx", "fp32c")
    I32 = ctx.dwba_intensity(ff32, fc=fcd)
    torch.cuda.synchronize()
    pix = pixel_errors()
    pix.add(I32, I)
    rel32 = float(torch.linalg.vector_norm(I32 - I) / torch.linalg.vector_norm(I))
    # parity on every 8th row (the oracle's exact method is slow), with the same degenerate-case
    # threshold as the GPU default (the reference GPU path's 1e-13)
    rows = np.arange(0, 4 * 256, 8)
    gs = O.Grid(256, len(rows), np.repeat(g.qx[:256][None], len(rows), 0).ravel(),
                np.repeat(g.qy[:256][None], len(rows), 0).ravel(), g.qz.reshape(4 * 256, 256)[rows].ravel(),
                g.alpha_f)
    E_ref = O.ff_exact(tri, gs, tiny=1e-13)
    E_gpu = ctx.as_complex(ffe).cpu().numpy().reshape(4 * 256, 256)[rows].ravel()
    rel_exact = float(np.linalg.norm(E_gpu - E_ref) / np.linalg.norm(E_ref))
    t0 = time.perf_counter()
    F_ref = O.ff_approx(O.tri9_to_rec8(tri), g)
    I_ref = O.dwba_intensity(g.npix, 4, fc, None, F_ref)
    t_cpu = time.perf_counter() - t0
    got = I.cpu().numpy()
    pairs = 4 * 256 * 256 * tri.shape[0]
    ctx.close()
    ka_ms = k_approx["main_ms"] / max(k_approx["main_launches"], 1)
    ke_ms = k_exact["main_ms"] / max(k_exact["main_launches"], 1)
    ke32_ms = k_exact32["main_ms"] / max(k_exact32["main_launches"], 1)
    fpp32, nt_mix = exact32_flop_per_pair(g, tri)
    fpp32 = float("nan") if fpp32 is None else fpp32
    # FP64 roofline of the form-factor kernels: algorithmic flop per pair (approximate on this product
    # grid: 16, two complex multiply-adds on the FP64 matrix cores; exact about 200 on the VALU;
    # transcendentals not counted) over the kernel's own time, against 78.6 TFLOP/s (the FP64 peak of
    # both the matrix cores and the vector units)
    roof = lambda fpp, ms, bound: {"bound": bound, "achieved": fpp * pairs / (ms * 1e-3) / 1e12,
                                   "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                                   "frac": fpp * pairs / (ms * 1e-3) / PEAK_FP64, "flop_per_pair": fpp,
                                   "avg_launch_ms": ms}
    return {"gpu_ms": 1e3 * t_gpu, "cpu_ms": 1e3 * t_cpu, "gpu_pairs_per_s": pairs / t_gpu,
            "cpu_pairs_per_s": pairs / t_cpu, "cpu_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "relL2_intensity_gpu_vs_oracle": float(np.linalg.norm(got - I_ref) / np.linalg.norm(I_ref)),
            "roofline_approx_kernel": dict(roof(16, ka_ms, "fp64-mfma"), kernel="k_ff_product64"),
            "exact_method": {"gpu_ms": 1e3 * t_exact, "gpu_pairs_per_s": pairs / t_exact,
                             "relL2_vs_oracle_sample": rel_exact, "roofline": roof(200, ke_ms, "fp64-valu"),
                             "fp32c": {"gpu_ms": 1e3 * t_exact32, "gpu_pairs_per_s": pairs / t_exact32,
                                       "speedup_vs_fp64": t_exact / t_exact32, "relL2_vs_fp64": rel_e32,
                                       "roofline": dict(roof(fpp32, ke32_ms, "fp32-valu"),
                                                        peak=PEAK_FP32 / 1e12,
                                                        frac=fpp32 * pairs / (ke32_ms * 1e-3) / PEAK_FP32,
                                                        achieved=fpp32 * pairs / (ke32_ms * 1e-3) / 1e12,
                                                        kernel="k_ff_exact32", term_mix=nt_mix)}},
            "fp32c": {"relL2_intensity_vs_fp64": rel32, "pixel_error_vs_fp64": pix.result()},
            "workload": "pyramid 2048 triangles, 256^2 DWBA (4 branches), FP64 approximate FF + DWBA; "
                        "exact-triangle FF of the same mesh and grid"}


def ewald_flop_per_pair(orders):
    """Matrix-core flop per pair the Ewald-grid kernel executes with series orders 0..K: split-fp16
    orders 0 and 1 take 48 (two complex multiply-adds x 4 real products x 3 cross terms), single-fp16
    orders from 2 up take 16."""
    return 48 * min(orders + 1, 2) + 16 * max(orders - 1, 0)


def bench_ewald(nshapes=64, reps=2):
    """The 64 seeded convex polyhedra of the batched workload on the reference's own q-grids: each on
    QGrid::create's q-space grid of its side (qy [-2, 2], qz [0, 2] 1/nm, 10 keV, alpha_i 0.2 deg) with
    the four DWBA branches of a particle in vacuum, FP32C approximate method, one call per shape. The
    Ewald-grid matrix-core kernel against the per-pair FP32C kernel on the same calls, both against
    the FP64 per-pair kernel (intensity relL2 and per-pixel errors)."""
    import torch
    from hipgisaxs_amd import _native as N
    from hipgisaxs_amd.api import Context

    k0 = 2 * math.pi / (1.23984 / 10.0)
    ai = math.radians(0.2)
    shapes = make_workload(nshapes)
    ctxs = {}
    for side in sorted({s["side"] for s in shapes}):
        c = Context(0)
        c.qgrid_qspace(side, side, -2.0, 0.0, 2.0, 2.0, ai, k0)
        c.qz_extended(k0, ai, 0j)
        ctxs[side] = c
    pairs = 0.0
    orders = {}
    flop_exec = 0.0
    for s in shapes:
        c = ctxs[s["side"]]
        s["mid"] = c.upload_mesh(s["tri"], "exact9")
        s["ntri"] = c.mesh_info(s["mid"])["ntri"]
        path, k = c.ff_numeric_path(s["mid"], "approx", "fp32c")
        orders[f"{path}:{k}"] = orders.get(f"{path}:{k}", 0) + 1
        p = float(s["ntri"]) * c.info()["nq"]
        pairs += p
        flop_exec += p * (ewald_flop_per_pair(k) if path == "ewald" else float("nan"))
        s["out"] = {m: torch.empty(2 * c.info()["nq"], dtype=torch.float64, device="cuda:0")
                    for m in ("ewald", "pair", "fp64")}

    def run(kind, flags, prec, n):
        for c in ctxs.values():
            c.profile_begin()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            for s in shapes:
                ctxs[s["side"]].ff_numeric(s["mid"], "approx", prec, flags=flags, out=s["out"][kind])
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / n
        prof = [c.profile_end() for c in ctxs.values()]
        kms = sum(p["main_ms"] for p in prof) / n
        return t, kms

    sched = {}
    for pipe in (1, 0):  # the pipelined schedule, then the default one (whose outputs stay)
        for c in ctxs.values():
            c.lib.hgx_tune_ewald(c.h, pipe)
        run("ewald", 0, "fp32c", 1)  # warm-up (first launches)
        sched["pipelined" if pipe else "default"] = run("ewald", 0, "fp32c", reps)
    t_ew, k_ew = sched["default"]
    t_pp, k_pp = run("pair", N.FLAG_NO_MFMA, "fp32c", reps)
    t_64, k_64 = run("fp64", 0, "fp64", 1)
    adj = bench_adjoint_any_grid(ctxs, shapes[::4])
    # the same 64 calls as one batch (meshes in one context, each item on its shape's grid), replayed
    # from its hipGraph: one launch per series order plus the finishing launch
    from hipgisaxs_amd.api import Batch

    main = Context(0)
    items = []
    for s in shapes:
        s["out"]["batch"] = torch.empty_like(s["out"]["ewald"])
        items.append((ctxs[s["side"]], main.upload_mesh(s["tri"], "exact9"), None, s["out"]["batch"]))
    bat = Batch(main, items, precision="fp32c")
    bat.run(use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bat.run(use_graph=True)
    torch.cuda.synchronize()
    t_bat = (time.perf_counter() - t0) / reps
    bnum = sum(float(torch.linalg.vector_norm(s["out"]["batch"] - s["out"]["ewald"])) ** 2 for s in shapes)
    bden = sum(float(torch.linalg.vector_norm(s["out"]["ewald"])) ** 2 for s in shapes)
    bat.release()
    main.close()
    pix = pixel_errors()
    worst = 0.0
    for s in shapes:
        I64, Iew = intensity(s["out"]["fp64"]), intensity(s["out"]["ewald"])
        pix.add(Iew, I64)
        worst = max(worst, float(torch.linalg.vector_norm(Iew - I64) / torch.linalg.vector_norm(I64)))
    num = sum(float(torch.linalg.vector_norm(intensity(s["out"]["ewald"]) - intensity(s["out"]["fp64"]))) ** 2
              for s in shapes)
    den = sum(float(torch.linalg.vector_norm(intensity(s["out"]["fp64"]))) ** 2 for s in shapes)
    for c in ctxs.values():
        c.close()
    ok = not math.isnan(flop_exec)
    return {"pairs": pairs, "series_orders": orders,
            "ewald": {"value": pairs / t_ew, "ms": 1e3 * t_ew, "kernel_ms": k_ew, "kernel": "k_ff_ewald16",
                      "roofline": {"bound": "mfma", "achieved": 16 * pairs / (k_ew * 1e-3) / 1e12,
                                   "peak": PEAK_FP16 / 1e12, "unit": "TFLOP/s",
                                   "frac": 16 * pairs / (k_ew * 1e-3) / PEAK_FP16, "flop_per_pair": 16,
                                   "executed_tflops": flop_exec / (k_ew * 1e-3) / 1e12 if ok else None}},
            "ewald_pipelined": {"value": pairs / sched["pipelined"][0], "kernel_ms": sched["pipelined"][1]},
            "ewald_batch_graph": {"value": pairs / t_bat, "ms": 1e3 * t_bat,
                                  "relL2_vs_single_calls": math.sqrt(bnum / bden)},
            "fp32c_per_pair": {"value": pairs / t_pp, "ms": 1e3 * t_pp, "kernel_ms": k_pp},
            "fp64_per_pair": {"value": pairs / t_64, "ms": 1e3 * t_64, "kernel_ms": k_64},
            "speedup_vs_fp32c_per_pair": t_pp / t_ew, "speedup_vs_fp64_per_pair": t_64 / t_ew,
            "relL2_intensity_vs_fp64": math.sqrt(num / den), "relL2_worst_shape": worst,
            "pixel_error_vs_fp64": pix.result(), "record_adjoint": adj,
            "workload": f"{nshapes} convex polyhedra of the batched workload, each on QGrid::create's q-space "
                        "grid of its side (128-1024, qy [-2, 2], qz [0, 2], 10 keV, 0.2 deg) with 4 DWBA branches, "
                        "FP32C approximate method, one call per shape"}


def bench_adjoint_any_grid(ctxs, shapes):
    """Record adjoint of L = Re sum_q W(q) F(q) (hgx_ff_numeric_adjoint) for the given shapes on their
    q-space grids with seeded random W: FP32C (the VALU kernel on these non-product grids) against FP64,
    rate in (q-point, triangle) pairs per second and the gradients' relL2."""
    import torch

    gen = torch.Generator(device="cuda:0").manual_seed(7)
    W = {side: torch.randn(2 * c.info()["nq"], dtype=torch.float64, device="cuda:0", generator=gen)
         for side, c in ctxs.items()}
    pairs = sum(float(s["ntri"]) * ctxs[s["side"]].info()["nq"] for s in shapes)

    def run(prec):
        for c in ctxs.values():
            c.profile_begin()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grads = [ctxs[s["side"]].ff_numeric_adjoint(s["mid"], W[s["side"]], prec) for s in shapes]
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        kms = sum(c.profile_end()["main_ms"] for c in ctxs.values())
        return t, kms, grads

    run("fp32c")  # warm-up (first launches, workspace)
    t32, k32, g32 = run("fp32c")
    t64, k64, g64 = run("fp64")
    num = sum(float(torch.linalg.vector_norm(a - b)) ** 2 for a, b in zip(g32, g64))
    den = sum(float(torch.linalg.vector_norm(b)) ** 2 for b in g64)
    return {"shapes": len(shapes), "pairs": pairs,
            "fp32c": {"value": pairs / t32, "ms": 1e3 * t32, "kernel_ms": k32, "kernel": "k_adjoint32"},
            "fp64": {"value": pairs / t64, "ms": 1e3 * t64, "kernel_ms": k64, "kernel": "k_adjoint64"},
            "speedup_fp32c_vs_fp64": t64 / t32, "relL2_fp32c_vs_fp64": math.sqrt(num / den)}


# Algorithmic work of the fused ensemble kernels as implemented (add/mul 1 flop, FMA 2; sin/cos/exp
# counted apart as T). Per (pixel, grain) of a lattice whose in-plane directions and grain centre give
# branch-independent factors (rotations about z): four in-plane phases (16), two Dirichlet kernels and
# the centre phase shared by the four branches (9 + 6 T), per branch the out-of-plane lattice phase,
# its Dirichlet kernel, the product with the centre phase and the complex multiply-add with the
# hoisted cell value (16 + 2 T, x4), weight and |amp|^2 (5): 94 flop + 14 T. A separable box cell adds
# its in-plane factor, rotated q and location phase per grain: 28 flop + 4 T. Per (pixel, branch) the
# hoisted cell (its nodes x ~55-60 flop) and the branch factor (8).
ENS_GRAIN_FLOP = 94
ENS_SEP_FLOP = 28


def ensemble_flop(npix, nbranch, cell_per_branch, ngrains, per_grain=ENS_GRAIN_FLOP):
    return npix * (nbranch * (cell_per_branch + 8) + ngrains * per_grain)


def bench_lattice(reps=5):
    """Spheres R ~ N(5, 0.5) nm (21 nodes on [3, 7]) on bcc a = 20 nm, 10x10x5, 8 azimuths, DW 0.5 nm,
    1024^2 detector (2 m, 172 um), 12 keV, alpha_i = 0.15 deg, DWBA over Si: fused kernel, both paths."""
    import torch
    from hipgisaxs_amd.api import Context, lattice_intensity, param_distribution

    k0 = 2 * math.pi / (1.23984 / 12.0)
    ai = math.radians(0.15)
    ctx = Context(0)
    _det_grid(ctx, 1024, 0.172, ai, k0)
    fc = ctx.fresnel([3.39e-6], [3.55e-8], [0.0], 0, k0, ai)
    v, w = param_distribution(3.0, 7.0, 21, 2, 5.0, 0.5)
    dims = [(v, w / w.sum())]
    a = 20.0
    la, lb, lc = (a, 0, 0), (0, a, 0), (a / 2, a / 2, a / 2)
    rots = np.stack([np.array([[math.cos(p), -math.sin(p), 0], [math.sin(p), math.cos(p), 0], [0, 0, 1]])
                     for p in (j * math.pi / 4 for j in range(8))])
    dn2 = -(2 * 6.53e-6 + 2j * 1.6e-8)
    out = {}
    res = {}
    for prec in ("fp32c", "fp64"):
        I = torch.empty(1024 * 1024, dtype=torch.float64, device="cuda:0")
        kw = dict(dw_sigma=0.5, dn2=dn2, precision=prec, fc=fc, out=I)
        lattice_intensity(ctx, "sphere", dims, la, lb, lc, (10, 10, 5), rots, None, **kw)
        torch.cuda.synchronize()
        ctx.profile_begin()
        t0 = time.perf_counter()
        for _ in range(reps):
            lattice_intensity(ctx, "sphere", dims, la, lb, lc, (10, 10, 5), rots, None, **kw)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        prof = ctx.profile_end()
        samples = 4 * 1024 * 1024 * 8 * v.size
        out[prec] = {"ms": 1e3 * wall, "kernel_ms": prof["main_ms"] / reps, "q_pt_samples_per_s": samples / wall}
        res[prec] = I
    d = torch.linalg.vector_norm(res["fp32c"] - res["fp64"]) / torch.linalg.vector_norm(res["fp64"])
    out["relL2_fp32c_vs_fp64"] = float(d)
    pix = pixel_errors()
    pix.add(res["fp32c"], res["fp64"])
    out["pixel_error_fp32c_vs_fp64"] = pix.result()
    # VALU roofline on the work the kernels do (ensemble_flop): the 8 azimuths rotate about z, so the
    # sphere form factor (about 55 flop per node) is evaluated once per (pixel, branch) and the grain
    # loop runs the structure factor alone; transcendentals not counted
    flop_step = ensemble_flop(1024 * 1024, 4, v.size * 55, 8)
    for prec, peak in (("fp32c", PEAK_FP32), ("fp64", PEAK_FP64)):
        a = flop_step / (out[prec]["kernel_ms"] * 1e-3)
        out[prec]["roofline"] = {"bound": "valu", "achieved": a / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                                 "frac": a / peak}
    out["speedup_fp32c_vs_fp64"] = out["fp32c"]["q_pt_samples_per_s"] / out["fp64"]["q_pt_samples_per_s"]
    out["workload"] = "spheres N(5,0.5) 21 nodes, bcc a=20 10x10x5, 8 azimuths, DW 0.5, 1024^2 DWBA, fused"
    ctx.close()
    return out


def bench_fit(evals=20):
    """Cylinders R=4, H=12 on hcp a=15, 20x20x1, 512^2 detector (2 m, 344 um), 12 keV, 0.15 deg, DWBA;
    target = Poisson(1e4 I(theta*)/max I) (seeded); chi^2 and its 6-parameter gradient."""
    import torch
    from hipgisaxs_amd.api import Context, FitPlan

    k0 = 2 * math.pi / (1.23984 / 12.0)
    ai = math.radians(0.15)
    ctx = Context(0)
    _det_grid(ctx, 512, 0.344, ai, k0)
    fc = ctx.fresnel([3.39e-6], [3.55e-8], [0.0], 0, k0, ai)
    ua, ub, uc = (1.0, 0, 0), (0.5, 0.866, 0), (0.5, 0.2887, 0.8165)
    theta_true = np.array([4.0, 12.0, 15.0, 0.0, 0.5, 1.0])
    I = torch.empty(512 * 512, dtype=torch.float64, device="cuda:0")
    zero = torch.zeros_like(I)
    plan0 = FitPlan(ctx, "cylinder", ua, ub, uc, (20, 20, 1), zero, fc=fc, dn2=-(2 * 6.53e-6 + 2j * 1.6e-8))
    plan0.fwd(theta_true, I_model=I)
    rng = np.random.default_rng(SEED + 4)
    Ih = I.cpu().numpy()
    obs = torch.from_numpy(rng.poisson(1e4 * Ih / Ih.max()).astype(np.float64) * (Ih.max() / 1e4)).cuda()
    plan = FitPlan(ctx, "cylinder", ua, ub, uc, (20, 20, 1), obs, fc=fc, dn2=-(2 * 6.53e-6 + 2j * 1.6e-8))
    theta = theta_true * np.array([1.02, 0.98, 1.01, 1.0, 1.1, 0.95]) + np.array([0, 0, 0, 0.01, 0, 0])
    plan.bwd(theta)
    t0 = time.perf_counter()
    for _ in range(evals):
        plan.fwd(theta)
    t_fwd = (time.perf_counter() - t0) / evals
    t0 = time.perf_counter()
    for _ in range(evals):
        chi2, grad = plan.bwd(theta)
    t_bwd = (time.perf_counter() - t0) / evals
    # a particle-swarm generation (64 parameter vectors) through the population entry point
    pop = theta_true * (1 + 0.02 * rng.standard_normal((64, 6)))
    pop[:, 3] = 0.01 * rng.standard_normal(64)
    plan.fwd_many(pop)
    t0 = time.perf_counter()
    plan.fwd_many(pop)
    t_pop = time.perf_counter() - t0
    # bounded L-BFGS from a start inside the true parameters' basin, 100 iterations
    lo = np.array([2.0, 6.0, 10.0, -0.1, 0.0, 0.1])
    hi = np.array([6.0, 20.0, 20.0, 0.1, 2.0, 10.0])
    start = theta_true * np.array([1.01, 0.99, 1.003, 1.0, 1.05, 0.97])
    start[3] = 0.005
    f_true = plan.fwd(theta_true)
    t0 = time.perf_counter()
    th, f_fit, its = plan.lbfgs(start, lo=lo, hi=hi, iterations=100, tol=1e-10)
    t_fit = time.perf_counter() - t0
    ctx.close()
    return {"chi2_grad_evals_per_s": 1.0 / t_bwd, "ms_fwd": 1e3 * t_fwd, "ms_grad": 1e3 * t_bwd,
            "grad_over_fwd": t_bwd / t_fwd, "fd_equivalent_fwds": 13, "chi2": chi2,
            "population_64_ms": 1e3 * t_pop, "population_evals_per_s": 64 / t_pop,
            "lbfgs": {"iterations": its, "s": t_fit, "objective": f_fit, "objective_at_truth": f_true,
                      "theta": [float(x) for x in th]},
            "workload": "cylinders R4 H12, hcp a=15 20x20x1, 512^2 DWBA, 6 parameters, FP64 adjoint"}


def bench_grating(ncand=256, iters=1):
    """Trapezoidal grating lines (top 20 nm, sidewall 82 deg, height 40 nm; pitch 100 nm, 200 periods,
    length 100 nm) meshed to 106 912 triangles, 1024^2 q-space grid qy [-2, 2], qz [0, 2] with four
    DWBA branches over Si, 10 keV, 0.2 deg; seeded Poisson target at 1e4 peak counts. ncand candidates
    (perturbed starts, 256 as configured) run forward + adjoint + L-BFGS update per hipGraph step; the
    200-iteration run's time is the measured per-iteration time of all candidates times 200."""
    import torch
    from hipgisaxs_amd.api import Context, GratingPlan

    k0 = 2 * math.pi / (1.23984 / 10.0)
    ai = math.radians(0.2)
    ctx = Context(0)
    ctx.qgrid_explicit(1024, 1024, -2.0, 0.0, 2.0, 2.0, ai, k0)
    ctx.qz_extended(k0, ai, 0j)
    fc = ctx.fresnel([4.88e-6], [7.37e-8], [0.0], 0, k0, ai)
    model = dict(length=100.0, pitch=100.0, periods=200, n_bottom=64, n_side=40, n_top=64, n_length=256,
                 dn2=2e-5 + 1e-7j)
    truth = np.array([20.0, math.radians(82.0), 40.0, 1.0])
    npix = 1024 * 1024
    zero = torch.zeros(npix, dtype=torch.float64, device="cuda:0")
    p1 = GratingPlan(ctx, 1, zero, fc=fc, **model)
    p1.set_theta(truth[None, :])
    I = p1.intensity(0).cpu().numpy()
    p1.release()
    truth[3] = 1e4 / I.max()
    rng = np.random.default_rng(SEED + 5)
    obs = torch.from_numpy(rng.poisson(1e4 * I / I.max()).astype(np.float64)).cuda()
    starts = truth * (1.0 + rng.uniform(-0.08, 0.08, size=(ncand, 4)))
    plan = GratingPlan(ctx, ncand, obs, fc=fc, **model)
    plan.set_theta(starts)
    plan.eval(use_graph=True)  # capture + first evaluation
    torch.cuda.synchronize()
    _, chi0, _ = plan.get()
    plan.lbfgs(1)  # capture of the L-BFGS step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.lbfgs(iters)
    torch.cuda.synchronize()
    t_iter = (time.perf_counter() - t0) / iters
    _, chi1, _ = plan.get()
    # one forward alone (mesh + FP32C form factor + intensity) for the gradient / forward cost ratio
    Ibuf = torch.empty(npix, dtype=torch.float64, device="cuda:0")
    plan.intensity(0, out=Ibuf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(4):
        plan.intensity(k, out=Ibuf)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / 4
    plan.release()
    ctx.close()
    nslots = (256 // 8 * 208 // 8 + 2 * 208 // 16) * 128
    pairs = 4 * npix * nslots  # forward (q-point x triangle slot) pairs per candidate
    # matrix work per pair: forward 48 flop (two complex products, 4 real products each, as 3 split-fp16
    # products of 2 flop); adjoint 72 flop (three passes f = 1, qz, qz^2 of one complex product each)
    mfma_flop = (48 + 72) * pairs * ncand
    return {"grad_evals_per_s": ncand / t_iter, "ms_per_iteration": 1e3 * t_iter, "candidates": ncand,
            "roofline": {"bound": "mfma", "achieved": mfma_flop / t_iter / 1e12, "peak": PEAK_FP16 / 1e12,
                         "unit": "TFLOP/s", "frac": mfma_flop / t_iter / PEAK_FP16,
                         "note": "whole L-BFGS iteration (forward, pixel pass, adjoint, update) over its wall time"},
            "ms_forward_per_candidate": 1e3 * t_fwd, "grad_over_fwd": t_iter / ncand / t_fwd,
            "forward_pairs_per_candidate": pairs,
            "fwd_plus_adjoint_pair_rate": ncand * pairs / t_iter,
            "s_for_200_iterations": t_iter * 200,
            "chi2_median_start": float(np.median(chi0)), "chi2_median_after": float(np.median(chi1)),
            "iterations_timed": iters,
            "workload": "trapezoid lines 106912 triangles, pitch 100 x 200, 1024^2 DWBA (4 branches), "
                        "FP32C forward + adjoint + L-BFGS per hipGraph step"}


HIG_GOLD = Path(__file__).resolve().parent / "tests" / "golden" / "hig"
# the reference's example inputs (examples/01-03, 07) at their own detector sizes and grain counts,
# rebuilt from the committed fixtures of the same models (which use smaller detectors and fewer grains)
HIG_EXAMPLES = {
    "01-pyramid": ("pyramid_dwba", {"resolution = [60 50]": "resolution = [300 250]"}),
    "02-spheres-bcc": ("spheres_bcc", {"resolution = [60 30]": "resolution = [600 300]",
                                       "maxgrains = [4 4 1]": "maxgrains = [{g} {g} 1]"}),
    "03-cylinders-hcp": ("cylinders_hcp", {"resolution = [48 32]": "resolution = [600 300]",
                                           "maxgrains = [3 3 1]": "maxgrains = [{g} {g} 1]"}),
    "07-gratings": ("box_grating", {"resolution = [50 30]": "resolution = [600 300]",
                                    "maxgrains = [3 3 1]": "maxgrains = [{g} {g} 1]"}),
}


def _hig_text(name, grains):
    fixture, subs = HIG_EXAMPLES[name]
    text = (HIG_GOLD / f"{fixture}.hig").read_text()
    for a, b in subs.items():
        assert a in text, (fixture, a)
        text = text.replace(a, b.format(g=grains))
    return text


def bench_hig_examples(reps=3, cpu=True):
    """The .hig driver on the reference's four example models at full size (600 x 300 or 300 x 250
    detector, 625 grains for the lattice examples): wall time per image of compute_gisaxs, FP64, and
    the CPU oracle of the same loop timed on 2 x 2 grains (extrapolated per grain)."""
    from hipgisaxs_amd import hig as H
    import oracle.hig as OH

    out = {}
    for name in HIG_EXAMPLES:
        text = _hig_text(name, 25)
        res = {}
        for prec in ("fp64", "fp32c"):
            sim = H.HigSimulation(text=text, precision=prec)
            sim.run(0)  # grid, buffers, first launches
            t0 = time.perf_counter()
            for _ in range(reps):
                img = sim.run(0)
            wall = (time.perf_counter() - t0) / reps
            c = sim.last_counts()
            res[prec] = {"ms_per_image": 1e3 * wall, "images_per_s": 1.0 / wall,
                         "fused_grains": c["fused_grains"], "per_grain_calls": c["grains"]}
            sim.close()
            # kernel times of one more run with HIP events around every launch
            sim = H.HigSimulation(text=text, precision=prec, flags=H.PROFILE)
            sim.run(0)
            sim.run(0)
            res[prec]["kernels"] = sim.kernel_ms()
            sim.close()
        ng_ = 625 if name != "01-pyramid" else 1
        if name.startswith(("02", "03", "07")):
            # fused ensemble kernels, VALU roofline on the work they do (ensemble_flop): the grains
            # rotate about z, so the sphere (about 55 flop) and cylinder (60) form factors are evaluated
            # once per (pixel, branch) and the grain loop runs the structure factor; the boxes of 07
            # split into a vertical factor per branch (about 12 flop) and an in-plane one per grain
            ffl = {"02": 55, "03": 60, "07": 12}[name[:2]]
            flop = ensemble_flop(img.size, 4, ffl, ng_, ENS_GRAIN_FLOP + (ENS_SEP_FLOP if name.startswith("07") else 0))
            for prec, peak in (("fp64", PEAK_FP64), ("fp32c", PEAK_FP32)):
                a_ = flop / (res[prec]["kernels"]["main_ms"] * 1e-3)
                res[prec]["roofline"] = {"bound": "valu", "achieved": a_ / 1e12, "peak": peak / 1e12,
                                         "unit": "TFLOP/s", "frac": a_ / peak}
        ng = 625 if name != "01-pyramid" else 1
        res["grains"] = ng
        res["pixels"] = int(img.size)
        if cpu:
            small = _hig_text(name, 2)
            t0 = time.perf_counter()
            OH.run(small)
            t_small = time.perf_counter() - t0
            per_grain = t_small / (4 if ng > 1 else 1)
            res["cpu_baseline"] = {"kind": "port", "sample": "2 x 2 grains" if ng > 1 else "1 grain", "s": t_small,
                                   "ms_per_image_extrapolated": 1e3 * per_grain * ng,
                                   "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
            res["speedup_fp64_vs_cpu"] = res["cpu_baseline"]["ms_per_image_extrapolated"] / res["fp64"]["ms_per_image"]
        out[name] = res
        print(f"[bench] hig {name} {res['fp64']['ms_per_image']:.2f} ms", file=sys.stderr, flush=True)
    return out


def bench_hig_fit(cpu=True):
    """Fitting a .hig model (tests/golden/hig/fit_cylinders.hig at a 600 x 300 detector): objective
    evaluations per second of the particle-swarm driver (16 particles x 4 generations, each evaluation
    = update_params + compute_gisaxs + the sqrt_unit_l2 distance on the device), with each generation
    read back once and particle by particle; the gradient's tangent pass against the 2P + 1 objective
    evaluations of central differences; LMVM on tangent passes and on central differences; and the
    CPU oracle's objective timed on one evaluation of the same model."""
    import tempfile

    from hipgisaxs_amd import hig as H
    import oracle.hig as OH

    text = (Path(ROOT) / "tests" / "golden" / "hig" / "fit_cylinders.hig").read_text()
    text = text.replace("resolution = [40 24]", "resolution = [600 300]")
    text = text.replace("pvalue = 6 }", "pvalue = 16 }")
    with tempfile.TemporaryDirectory() as d:
        sim = H.HigSimulation(text=text, base_dir=d)
        for var, v in (("shape['cyl']:param['radius']:min", 5.0), ("structure['hcp']:grain:scaling", 50.0)):
            sim.update_param(var, v)
        H.HigSimulation.save(sim.compute(), str(Path(d) / "fit_cylinders_ref.out"))
        sim.close()
        sim = H.HigSimulation(text=text, base_dir=d)
        sim.fit_init()
        sim.objective([4.2, 48.5])  # first launches
        t0 = time.perf_counter()
        out = sim.fit(algorithm=0, seed=1)
        wall = time.perf_counter() - t0
        res = {"pixels": 600 * 300, "pso_evals": out["evals"], "pso_s": wall, "evals_per_s": out["evals"] / wall,
               "ms_per_eval": 1e3 * wall / out["evals"], "best": [float(x) for x in out["x"]]}
        sim.close()
        seq = H.HigSimulation(text=text, base_dir=d, flags=H.FIT_SEQUENTIAL)
        seq.fit_init()
        seq.objective([4.2, 48.5])
        t0 = time.perf_counter()
        out_s = seq.fit(algorithm=0, seed=1)
        res["pso_particle_by_particle_s"] = time.perf_counter() - t0
        res["pso_same_result"] = bool(np.array_equal(out_s["x"], out["x"]) and out_s["value"] == out["value"])
        seq.close()
        # the gradient: one tangent pass against 2P + 1 objective evaluations
        sim = H.HigSimulation(text=text, base_dir=d)
        sim.fit_init()
        sim.gradient([4.2, 48.5])
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            sim.gradient([4.2, 48.5])
        t_grad = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            sim.objective([4.2, 48.5])
        t_obj = (time.perf_counter() - t0) / reps
        res["gradient"] = {"tangent_pass_ms": 1e3 * t_grad, "objective_ms": 1e3 * t_obj,
                           "central_differences_ms": 1e3 * 5 * t_obj, "tangent_over_objective": t_grad / t_obj}
        fine = text.replace("step = 0.5 }", "step = 0.02 }").replace("step = 2.5 }", "step = 0.1 }")
        for name, flags in (("lmvm_tangent", 0), ("lmvm_central_differences", H.FIT_FD_GRADIENT)):
            s2 = H.HigSimulation(text=fine, base_dir=d, flags=flags)
            s2.fit_init()
            t0 = time.perf_counter()
            o2 = s2.fit(algorithm=2)
            res[name] = {"s": time.perf_counter() - t0, "evals": o2["evals"], "value": o2["value"],
                         "x": [float(x) for x in o2["x"]]}
            s2.close()
        sim.close()
        if cpu:
            params, _, _ = OH.read_fitting(text)
            O = OH.read_ascii_image(str(Path(d) / "fit_cylinders_ref.out"))
            t0 = time.perf_counter()
            OH.objective(text, params, [4.2, 48.5], "sqrt_unit_l2", O)
            tc = time.perf_counter() - t0
            res["cpu_baseline"] = {"kind": "port", "sample": "one objective evaluation", "s": tc,
                                   "evals_per_s": 1.0 / tc,
                                   "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    print(f"[bench] hig fit {res['ms_per_eval']:.3f} ms per evaluation", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shapes", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the lattice and chi^2 measurements")
    ap.add_argument("--mode", choices=("batch", "per-shape"), default="batch",
                    help="batch: one launch for all shapes (hipGraph replay); per-shape: one call per shape")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the work-unit shape first (default: the library's default shape)")
    ap.add_argument("--no-tune", action="store_true", help="(default) skip the work-unit autotuning")
    ap.add_argument("--operands", choices=("fp32", "f16x2"), default=None,
                    help="product-grid operand precision when not autotuning (default: the library's)")
    ap.add_argument("--row-tiles", type=int, default=None, choices=(1, 2, 4),
                    help="split-fp16 product kernel: row tiles per wave when not autotuning")
    ap.add_argument("--flat", action="store_true", help="with --row-tiles 4: one-level accumulation")
    ap.add_argument("--no-share", action="store_true", help="with --row-tiles: row operands built per wave")
    ap.add_argument("--blocks-per-unit", type=int, default=None,
                    help="product kernel: triangle blocks per work unit when not autotuning")
    ap.add_argument("--kahan", action="store_true", help="per-pair Kahan accumulation in FP32C")
    ap.add_argument("--no-mfma", action="store_true",
                    help="per-pair VALU kernels for FP32C even on product grids (no matrix cores)")
    ap.add_argument("--hig-only", action="store_true",
                    help="only the .hig driver leg (the reference's example models), as one JSON object")
    ap.add_argument("--leg", choices=("pyramid", "lattice", "fit", "grating", "hig", "ewald"), default=None,
                    help="only that measurement leg, as one JSON object (for profiling one workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build is single-GPU); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from hipgisaxs_amd.api import Batch, Context

    if args.hig_only or args.leg == "hig":
        print(json.dumps({"hig_examples": bench_hig_examples(cpu=not args.no_cpu),
                          "hig_fit": bench_hig_fit(cpu=not args.no_cpu)}))
        return
    if args.leg is not None:
        legs = {"pyramid": bench_pyramid_fp64, "lattice": bench_lattice, "fit": bench_fit, "grating": bench_grating,
                "ewald": bench_ewald}
        print(json.dumps({args.leg: legs[args.leg]()}))
        return
    if args.copy_probe:
        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "copies": 3}))
        return

    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    shapes = make_workload(args.shapes)
    sides = sorted({s["side"] for s in shapes})
    grid_ctx = {}
    for side in sides:
        c = Context(0)
        c.qgrid_explicit(side, side, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
        grid_ctx[side] = c
    main_ctx = Context(0)  # holds every mesh and the batches; grids live in grid_ctx
    for s in shapes:
        s["mid"] = main_ctx.upload_mesh(s["tri"], "exact9")
        s["ntri"] = main_ctx.mesh_info(s["mid"])["ntri"]
        if args.mode == "per-shape":  # per-shape calls run in the context that holds the shape's grid
            s["mid_grid"] = grid_ctx[s["side"]].upload_mesh(s["tri"], "exact9")
        s["out32"] = torch.empty(2 * s["side"] ** 2, dtype=torch.float64, device=dev)
        s["out64"] = torch.empty(2 * s["side"] ** 2, dtype=torch.float64, device=dev)
    pairs = float(sum(s["ntri"] * s["side"] ** 2 for s in shapes))
    flags = (2 if args.kahan else 0) | (4 if args.no_mfma else 0)
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s: {len(shapes)} shapes, "
          f"{sum(s['ntri'] for s in shapes)} triangles, {pairs:.3e} pairs/step", file=sys.stderr, flush=True)

    items = {prec: [(grid_ctx[s["side"]], s["mid"], None, s["out32" if prec == "fp32c" else "out64"])
                    for s in shapes] for prec in ("fp32c", "fp64")}
    # every grid here is a product grid (qx = 0, qy per column, qz per row): FP32C runs on the
    # matrix cores unless --no-mfma
    product = all(c.is_product() for c in grid_ctx.values()) and not args.no_mfma
    tuned = {}
    from hipgisaxs_amd.api import tune_product
    from hipgisaxs_amd.autotune import PRODUCT_DEFAULT

    pcfg = tuple(PRODUCT_DEFAULT)
    if args.operands is not None:
        pcfg = (16, 3, "fp32") if args.operands == "fp32" else pcfg
    if args.row_tiles is not None:
        pcfg = (pcfg[0], 4, "f16x2", args.row_tiles, not args.flat, not args.no_share)
    if args.blocks_per_unit is not None:
        pcfg = (args.blocks_per_unit,) + pcfg[1:]
    tune_product(main_ctx, *pcfg)
    if args.mode == "batch" and args.tune and not args.no_tune:
        from hipgisaxs_amd.autotune import autotune_batch, autotune_product

        if product:
            best, times = autotune_product(main_ctx, items["fp32c"], flags)
            pcfg = tuple(best)
            tuned = {"product_config": "x".join(map(str, best)),
                     "candidates_ms": {"x".join(map(str, k)): round(v, 3) for k, v in times.items()}}
        else:
            best, times = autotune_batch(main_ctx, items["fp32c"], "fp32c", flags)
            tuned = {"qpt": best[0], "blocks_per_unit": best[1],
                     "candidates_ms": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in times.items()}}
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)

    def make_step(prec):
        if args.mode == "batch":
            b = Batch(main_ctx, items[prec], precision=prec, flags=flags if prec == "fp32c" else 0)
            return lambda graph: b.run(use_graph=graph), b
        def per_shape(graph):
            for s in shapes:
                grid_ctx[s["side"]].ff_numeric(s["mid_grid"], "approx", prec, flags=flags if prec == "fp32c" else 0,
                                               out=s["out32" if prec == "fp32c" else "out64"])
        return per_shape, None

    def timed(step, steps, warmup, graph):
        for _ in range(warmup):
            step(graph)
        torch.cuda.synchronize()
        for c in [main_ctx, *grid_ctx.values()]:
            c.profile_begin()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(graph)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = [c.profile_end() for c in [main_ctx, *grid_ctx.values()]]
        return wall, sum(p["main_ms"] for p in prof), sum(p["main_launches"] for p in prof), \
            sum(p["aux_ms"] for p in prof)

    step32, b32 = make_step("fp32c")
    use_graph = args.mode == "batch"
    # headline: the whole step replayed from the captured hipGraph
    wall32, _, _, _ = timed(step32, args.steps, args.warmup, use_graph)
    print(f"[bench] fp32c {args.steps} steps: {wall32:.3f}s", file=sys.stderr, flush=True)
    # roofline: the same kernels launched directly, so the hot kernel's own events are recorded
    _, k32_ms, n32, aux32 = timed(step32, args.steps, 0, False)
    # plain FP64 on the same tiling (the FP64 product-grid kernel on product grids), then the
    # per-pair FP64 kernels (HGX_FLAG_NO_MFMA) as a second baseline
    step64, b64 = make_step("fp64")
    wall64, k64_ms, n64, aux64 = timed(step64, args.fp64_steps, 1 if args.warmup else 0, use_graph)
    print(f"[bench] fp64 {args.fp64_steps} steps: {wall64:.3f}s", file=sys.stderr, flush=True)
    wall64p = k64p_ms = None
    if product and args.mode == "batch" and not args.no_extra:
        bp = Batch(main_ctx, items["fp64"], precision="fp64", flags=4)
        wall64p, k64p_ms, _, _ = timed(lambda g: bp.run(use_graph=g), 1, 1, True)
        bp.release()
        step64(use_graph)  # the same-tiling FP64 outputs again, as the reference for the accuracy below
        torch.cuda.synchronize()
        print(f"[bench] fp64 per pair 1 step: {wall64p:.3f}s", file=sys.stderr, flush=True)
    if b64 is not None:
        b64.release()

    # accuracy of the FP32C outputs against the plain FP64 GPU outputs (all shapes): relL2 of the
    # amplitudes, and per pixel of the intensity I = |F|^2
    num = den = 0.0
    worst = 0.0
    pix = pixel_errors()
    for s in shapes:
        a, b = s["out32"], s["out64"]
        d = float(torch.linalg.vector_norm(a - b)) ** 2
        n = float(torch.linalg.vector_norm(b)) ** 2
        num += d
        den += n
        worst = max(worst, math.sqrt(d / max(n, 1e-300)))
        pix.add(intensity(a), intensity(b))
    rel_fp64 = math.sqrt(num / den)

    value32 = pairs * args.steps / wall32
    value64 = pairs * args.fp64_steps / wall64
    avg_launch_ms = k32_ms / max(n32, 1)
    pairs_per_launch = pairs * args.steps / max(n32, 1)
    bpu_p, mults, operands = pcfg[:3]
    # algorithmic work: on product grids two complex multiply-adds per (q, triangle) pair (16 flop,
    # the factorized form's own count; the split-fp16 emulation and the table work are overhead), per
    # pair the direct form's 33 flop otherwise
    fpp = 16 if product else FLOP_PER_PAIR
    fpp_exec = flop_per_pair_product(mults, operands) if product else FLOP_PER_PAIR
    achieved = fpp * pairs_per_launch / (avg_launch_ms * 1e-3)
    achieved_exec = fpp_exec * pairs_per_launch / (avg_launch_ms * 1e-3)
    split16 = product and operands == "f16x2"
    peak = PEAK_FP16 if split16 else PEAK_FP32
    result = {
        "metric": METRIC,
        "value": value32,
        "unit": "q-pt·triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x2-split/fp32-acc" if split16 else "fp32",
        "data": "synthetic",
        "config": {
            "workload": "batched numeric FF: 64 seeded convex polyhedra (ellipsoid semi-axes U[3,15] nm), "
                        "500-50000 triangles log-uniform, own Born q-grids 128^2-1024^2 over "
                        "qy[-2,2] qz[0,2] nm^-1, approximate-triangle method",
            "shapes": len(shapes),
            "triangles": int(sum(s["ntri"] for s in shapes)),
            "pairs_per_step": pairs,
            "accumulation": "kahan" if args.kahan else "fp32-block/fp64-carry",
            "mode": args.mode + (" (hipGraph replay)" if use_graph else ""),
            "tuning": tuned,
        },
        "fp64_plain": {"value": value64, "ms_per_step": 1e3 * wall64 / args.fp64_steps,
                       "kernel_ms_per_step": k64_ms / args.fp64_steps,
                       "kernel": "k_ff_product64 (same tiling, FP64 matrix cores)" if product else
                                 ("k_batch64" if args.mode == "batch" else "k_ff_approx64")},
        "speedup_fp32c_vs_fp64": value32 / value64,
        "relL2_fp32c_vs_fp64_gpu": rel_fp64,
        "relL2_fp32c_vs_fp64_gpu_worst_shape": worst,
        "pixel_error_fp32c_vs_fp64_gpu": pix.result(),
        "roofline": {
            "bound": "mfma" if product else "valu",
            "kernel": (("k_ff_product16s" if len(pcfg) > 5 and pcfg[5] else "k_ff_product16") if split16
                       else "k_ff_product32") if product else
                      ("k_batch32" if args.mode == "batch" else "k_ff_approx32"),
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": None,
            "flop_per_pair": fpp,
            "executed": {"flop_per_pair": fpp_exec, "tflops": achieved_exec / 1e12, "frac": achieved_exec / peak},
            "avg_launch_ms": avg_launch_ms,
            "launches": n32,
            "aux_ms_per_step": aux32 / args.steps,
        },
    }
    if wall64p is not None:
        result["fp64_per_pair"] = {"value": pairs / wall64p, "ms_per_step": 1e3 * wall64p,
                                   "kernel": "k_batch64 (one q-point per lane, no matrix cores)"}
        result["speedup_fp32c_vs_fp64_per_pair"] = value32 * wall64p / pairs
    if product:
        rate = pairs_per_launch / (avg_launch_ms * 1e-3)
        # the same pairs counted in the direct per-pair form (33 flop): a rate of equivalent work
        result["roofline"]["direct_form_equivalent_tflops"] = FLOP_PER_PAIR * rate / 1e12
        if len(shapes) == 64 and not args.kahan and pcfg in PRODUCT_TRAFFIC_BYTES:
            kb, fb = PRODUCT_TRAFFIC_BYTES[pcfg]
            result["roofline"]["traffic"] = kb
            # the whole step: the finishing kernel reads every chunk's FP32 partials once more
            result["roofline"]["traffic_step"] = kb + fb if fb is not None else None
            result["roofline"]["algorithmic_bytes"] = ALGO_BYTES(shapes)
    if b32 is not None:
        result["config"]["work_units"] = b32.info()["units"]
    if product and not args.no_extra:
        # the same batch on the per-pair VALU kernels (no matrix cores), for comparison
        bv = Batch(main_ctx, items["fp32c"], precision="fp32c", flags=flags | 4)
        wall_v, kv_ms, nv, _ = timed(lambda g: bv.run(use_graph=g), 2, 1, True)
        result["fp32c_per_pair_valu"] = {"value": pairs * 2 / wall_v, "ms_per_step": 1e3 * wall_v / 2}
        bv.release()
        # the one-level accumulation variant (FP32 sums over a whole work unit): faster, less accurate
        from hipgisaxs_amd.autotune import PRODUCT_ONE_LEVEL

        tune_product(main_ctx, *PRODUCT_ONE_LEVEL)
        b1 = Batch(main_ctx, items["fp32c"], precision="fp32c", flags=flags)
        wall_1, k1_ms, n1, _ = timed(lambda g: b1.run(use_graph=g), args.steps, 1, True)
        b1.release()
        num1 = den1 = 0.0
        for s_ in shapes:
            num1 += float(torch.linalg.vector_norm(s_["out32"] - s_["out64"])) ** 2
            den1 += float(torch.linalg.vector_norm(s_["out64"])) ** 2
        result["fp32c_one_level"] = {"config": "x".join(map(str, PRODUCT_ONE_LEVEL)),
                                     "value": pairs * args.steps / wall_1, "ms_per_step": 1e3 * wall_1 / args.steps,
                                     "relL2_vs_fp64_gpu": math.sqrt(num1 / den1)}
        tune_product(main_ctx, *pcfg)
        step32(False)  # the headline outputs again, for the oracle comparison below
        torch.cuda.synchronize()
    if not args.no_extra:
        result["pyramid_fp64"] = bench_pyramid_fp64()
        print(f"[bench] pyramid done", file=sys.stderr, flush=True)
        result["lattice_spheres_bcc"] = bench_lattice()
        print(f"[bench] lattice done", file=sys.stderr, flush=True)
        result["chi2_cylinders_hcp"] = bench_fit()
        print(f"[bench] fit done", file=sys.stderr, flush=True)
        result["grating_fit"] = bench_grating()
        print(f"[bench] grating done", file=sys.stderr, flush=True)
        result["ewald_grids"] = bench_ewald(args.shapes)
        print(f"[bench] ewald grids done", file=sys.stderr, flush=True)
        result["hig_examples"] = bench_hig_examples(cpu=not args.no_cpu)
        result["hig_fit"] = bench_hig_fit(cpu=not args.no_cpu)
        print(f"[bench] hig examples done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cb, (s, rows, ref) = cpu_baseline(shapes)
        got = Context.as_complex(s["out32"]).cpu().numpy()[: rows * s["side"]]
        result["relL2_fp32c_vs_oracle_sample"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        got64 = Context.as_complex(s["out64"]).cpu().numpy()[: rows * s["side"]]
        result["relL2_fp64_vs_oracle_sample"] = float(np.linalg.norm(got64 - ref) / np.linalg.norm(ref))
        result["cpu_baseline"] = cb
    print(json.dumps(result))


if __name__ == "__main__":
    main()
