"""Autotuner for the launch shapes of the numeric form-factor kernels.

Times each candidate (q-points per thread, triangle blocks per work unit) on the caller's own batch
with HIP events (the library's launch timing) and keeps the fastest. The reference's counterpart is
the hyperblock-size search of its legacy CPU path (src/ff/cpu/ff_num_cpu.cpp:116-176).
"""
from __future__ import annotations

from .api import Batch, Context, tune_batch

DEFAULT_CANDIDATES = ((4, 8), (4, 16), (4, 32), (8, 16), (2, 16), (8, 32))


def autotune_batch(ctx: Context, items, precision: str = "fp32c", flags: int = 0,
                   candidates=DEFAULT_CANDIDATES, reps: int = 1):
    """Returns ((qpt, blocks_per_unit), {candidate: ms}) and leaves the best setting active."""
    import torch

    times = {}
    for qpt, bpu in candidates:
        tune_batch(qpt, bpu)
        b = Batch(ctx, items, precision=precision, flags=flags)
        b.run()
        torch.cuda.synchronize()
        ctx.profile_begin()
        for _ in range(reps):
            b.run()
        prof = ctx.profile_end()
        times[(qpt, bpu)] = (prof["main_ms"] + prof["aux_ms"]) / reps
        b.release()
    best = min(times, key=times.get)
    tune_batch(*best)
    return best, times
