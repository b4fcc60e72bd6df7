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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "q-pt·triangles/s + χ²-grad evals/s, FP32-comp vs plain FP64 GPU @ relL2≤1e-5"
SEED = 20261203
PEAK_FP32 = 157.3e12  # MI355X dense FP32 vector (packed) peak
PEAK_FP64 = 78.6e12
FLOP_PER_PAIR = 33  # approximate-triangle pair: q.n 10 + q.c 10 + phase 2 + scale 3 + cmul 6 + acc 2


def make_workload(nshapes=64, seed=SEED + 3):
    from hipgisaxs_amd import meshes

    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(nshapes):
        axes = rng.uniform(3.0, 15.0, 3)
        target = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        side = int(rng.choice([128, 256, 512, 1024]))
        tri = meshes.convex_polyhedron(rng, axes, target)
        shapes.append({"tri": tri, "side": side, "axes": axes})
    return shapes


def cpu_baseline(shapes, budget_s=12.0):
    """Times the FP64 OpenMP oracle on a bounded sample: one mid-size shape, a band of q-rows."""
    import oracle as O

    s = sorted(shapes, key=lambda d: d["tri"].shape[0])[len(shapes) // 2]
    rec = O.tri9_to_rec8(s["tri"])
    side = s["side"]
    g = O.qgrid_explicit(side, side, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
    g = O.Grid(side, side, g.qx, g.qy, g.qz.real + 0j, None)
    # calibrate on a few rows, then size the sample to about budget_s
    rows = 2
    while True:
        sub = O.Grid(side, rows, g.qx[: rows * side], g.qy[: rows * side], g.qz[: rows * side])
        t0 = time.perf_counter()
        O.ff_approx(rec, sub)
        dt = time.perf_counter() - t0
        if dt > 0.5 or rows >= side:
            break
        rows = min(side, rows * 4)
    rows = int(min(side, max(1, rows * budget_s / max(dt, 1e-3))))
    sub = O.Grid(side, rows, g.qx[: rows * side], g.qy[: rows * side], g.qz[: rows * side])
    t0 = time.perf_counter()
    out = O.ff_approx(rec, sub)
    dt = time.perf_counter() - t0
    pairs = rows * side * rec.shape[0]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": pairs / dt,
        "unit": "q-pt·triangles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 OpenMP oracle, approximate method, shape with {rec.shape[0]} triangles x "
                  f"{rows} rows of its {side}x{side} grid ({pairs:.3e} pairs, {dt:.1f} s)",
    }, (s, rows, out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shapes", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--kahan", action="store_true", help="per-pair Kahan accumulation in FP32C")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build is single-GPU); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from hipgisaxs_amd.api import Context

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
    ctxs = {}
    for side in sides:
        c = Context(0)
        c.qgrid_explicit(side, side, -2.0, 0.0, 2.0, 2.0, 0.0, 1e3)
        ctxs[side] = c
    for s in shapes:
        c = ctxs[s["side"]]
        s["mid"] = c.upload_mesh(s["tri"], "exact9")
        s["ntri"] = c.mesh_info(s["mid"])["ntri"]
        s["out32"] = torch.empty(2 * s["side"] ** 2, dtype=torch.float64, device=dev)
        s["out64"] = torch.empty(2 * s["side"] ** 2, dtype=torch.float64, device=dev)
    pairs = float(sum(s["ntri"] * s["side"] ** 2 for s in shapes))
    flags = 2 if args.kahan else 0
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s: {len(shapes)} shapes, "
          f"{sum(s['ntri'] for s in shapes)} triangles, {pairs:.3e} pairs/step", file=sys.stderr, flush=True)

    def step(prec):
        key = "out32" if prec == "fp32c" else "out64"
        for s in shapes:
            ctxs[s["side"]].ff_numeric(s["mid"], "approx", prec, flags=flags if prec == "fp32c" else 0, out=s[key])

    def timed(prec, steps, warmup):
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        for c in ctxs.values():
            c.profile_begin()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(prec)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = [c.profile_end() for c in ctxs.values()]
        main_ms = sum(p["main_ms"] for p in prof)
        main_n = sum(p["main_launches"] for p in prof)
        aux_ms = sum(p["aux_ms"] for p in prof)
        return wall, main_ms, main_n, aux_ms

    wall32, k32_ms, n32, aux32 = timed("fp32c", args.steps, args.warmup)
    print(f"[bench] fp32c {args.steps} steps: {wall32:.3f}s", file=sys.stderr, flush=True)
    wall64, k64_ms, n64, aux64 = timed("fp64", args.fp64_steps, 1 if args.warmup else 0)
    print(f"[bench] fp64 {args.fp64_steps} steps: {wall64:.3f}s", file=sys.stderr, flush=True)

    # accuracy of the FP32C outputs against the plain FP64 GPU outputs (all shapes)
    num = den = 0.0
    worst = 0.0
    for s in shapes:
        a, b = s["out32"], s["out64"]
        d = float(torch.linalg.vector_norm(a - b)) ** 2
        n = float(torch.linalg.vector_norm(b)) ** 2
        num += d
        den += n
        worst = max(worst, math.sqrt(d / max(n, 1e-300)))
    rel_fp64 = math.sqrt(num / den)

    value32 = pairs * args.steps / wall32
    value64 = pairs * args.fp64_steps / wall64
    avg_launch_ms = k32_ms / max(n32, 1)
    flops_per_launch = FLOP_PER_PAIR * pairs / len(shapes)
    achieved = flops_per_launch / (avg_launch_ms * 1e-3)
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
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "batched numeric FF: 64 seeded convex polyhedra (ellipsoid semi-axes U[3,15] nm), "
                        "500-50000 triangles log-uniform, own Born q-grids 128^2-1024^2 over "
                        "qy[-2,2] qz[0,2] nm^-1, approximate-triangle method",
            "shapes": len(shapes),
            "triangles": int(sum(s["ntri"] for s in shapes)),
            "pairs_per_step": pairs,
            "accumulation": "kahan" if args.kahan else "fp32-block/fp64-carry",
        },
        "fp64_plain": {"value": value64, "ms_per_step": 1e3 * wall64 / args.fp64_steps,
                       "kernel_ms_per_step": k64_ms / args.fp64_steps},
        "speedup_fp32c_vs_fp64": value32 / value64,
        "relL2_fp32c_vs_fp64_gpu": rel_fp64,
        "relL2_fp32c_vs_fp64_gpu_worst_shape": worst,
        "roofline": {
            "bound": "valu",
            "kernel": "k_ff_approx32",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": None,
            "flop_per_pair": FLOP_PER_PAIR,
            "avg_launch_ms": avg_launch_ms,
            "launches": n32,
            "aux_ms_per_step": aux32 / args.steps,
        },
    }
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
