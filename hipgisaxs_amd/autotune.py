"""Autotuner for the launch shapes of the batched numeric form-factor kernels.

Times each candidate on the caller's own batch with HIP events (the library's launch timing) and
keeps the fastest; a candidate replaces the default only when it is faster by more than `margin`,
so repeated runs settle on the same shape. Two kernel families are tuned:

* the per-pair VALU kernels: (q-points per thread, triangle blocks per work unit);
* the product-grid matrix-core kernel: (triangle blocks per work unit, real products per complex
  product: 3 or 4, operand precision: FP32 or split fp16, and for split fp16 the row tiles per wave
  and two-level accumulation).

The reference's counterpart is the hyperblock-size search of its legacy CPU path
(src/ff/cpu/ff_num_cpu.cpp:116-176).
"""
from __future__ import annotations

from .api import PRODUCT_DEFAULT, Batch, Context, tune_batch, tune_product

VALU_DEFAULT = (4, 16)
VALU_CANDIDATES = (VALU_DEFAULT, (4, 8), (4, 32), (8, 16), (2, 16), (8, 32))
# two-level accumulation only: the one-level variants (FP32 sums over a whole work unit) are about
# 10x less accurate and are measured separately
PRODUCT_CANDIDATES = (PRODUCT_DEFAULT, (16, 4, "f16x2", 2, True, True), (32, 4, "f16x2", 2, True, True),
                      (64, 4, "f16x2", 4, True, True), (32, 4, "f16x2", 4, True, True),
                      (64, 4, "f16x2", 2, True, False), (64, 4, "f16x2", 4, True, False),
                      (16, 4, "f16x2", 1, True, False), (16, 3, "fp32"))
PRODUCT_ONE_LEVEL = (64, 4, "f16x2", 4, False, True)


def _time_batch(ctx: Context, items, precision, flags, reps):
    import torch

    b = Batch(ctx, items, precision=precision, flags=flags)
    b.run()
    torch.cuda.synchronize()
    ctx.profile_begin()
    for _ in range(reps):
        b.run()
    prof = ctx.profile_end()
    b.release()
    return (prof["main_ms"] + prof["aux_ms"]) / reps


def _search(ctx, setter, default, candidates, run, margin):
    times = {}
    for cand in candidates:
        setter(ctx, *cand)
        times[cand] = run()
    best = min(times, key=times.get)
    if default in times and times[best] > (1.0 - margin) * times[default]:
        best = default
    setter(ctx, *best)
    return best, times


def autotune_batch(ctx: Context, items, precision: str = "fp32c", flags: int = 0,
                   candidates=VALU_CANDIDATES, reps: int = 1, margin: float = 0.02):
    """Per-pair kernels: returns ((qpt, blocks_per_unit), {candidate: ms}), leaves the best active
    in the context."""
    return _search(ctx, tune_batch, VALU_DEFAULT, candidates,
                   lambda: _time_batch(ctx, items, precision, flags, reps), margin)


def autotune_product(ctx: Context, items, flags: int = 0, candidates=PRODUCT_CANDIDATES, reps: int = 1,
                     margin: float = 0.02):
    """Product-grid kernel (FP32C): returns (best candidate tuple as taken by tune_product, {candidate: ms})."""
    return _search(ctx, tune_product, PRODUCT_DEFAULT, candidates,
                   lambda: _time_batch(ctx, items, "fp32c", flags, reps), margin)
