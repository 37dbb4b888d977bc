"""Launch-shape autotuner for the apply and assembly kernels.

Times each candidate with HIP events on the current stream (the stream the
library launches on) and keeps the fastest in the library context:
  apply:    CSR (lanes per row, nonzeros per lane, load width) x threads per block;
            sliced-ELL groups of 4 slots loaded ahead {2, 4}
  assembly: particles each wavefront assembles in sequence {1, 2, 4, 8, 16}
"""
from __future__ import annotations

import torch

from . import _lib


def _time(fn, reps):
    fn()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps


# (lanes per row, elements per lane, load width)
APPLY_SHAPES = ((4, 8, 4), (8, 4, 4), (8, 8, 4), (8, 12, 4), (16, 4, 4), (16, 8, 4), (32, 4, 4),
                (4, 16, 1), (8, 8, 1), (16, 4, 1), (32, 2, 1), (64, 1, 1), (8, 16, 1), (16, 8, 1))


def tune_apply(fn, shapes=APPLY_SHAPES, blocks=(256, 512), reps=10, device=0):
    """fn() runs one apply; returns ((G, epl, vec, block), ms) of the fastest shape."""
    ctx = _lib.context(device)
    best = None
    results = {}
    try:
        for g, e, v in shapes:
            for b in blocks:
                if b % g:
                    continue
                ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, g, e, v, b), "dcpse_set_apply_config")
                ms = _time(fn, reps)
                results[(g, e, v, b)] = ms
                if best is None or ms < best[1]:
                    best = ((g, e, v, b), ms)
    finally:
        cfg = best[0] if best else (0, 8, 4, 256)
        ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, *cfg), "dcpse_set_apply_config")
    return best, results


SELL_PIPELINES = ((2, 2), (4, 2), (1, 3), (2, 3))  # (groups per batch, stages)


def tune_sell(fn, pipelines=SELL_PIPELINES, reps=10, device=0):
    """fn() runs one sliced-ELL apply; returns ((groups, depth), ms) of the fastest."""
    ctx = _lib.context(device)
    best = None
    results = {}
    try:
        for gd in pipelines:
            ctx.check(ctx.lib.dcpse_set_sell_pipeline(ctx.h, *gd), "dcpse_set_sell_pipeline")
            ms = _time(fn, reps)
            results[gd] = ms
            if best is None or ms < best[1]:
                best = (gd, ms)
    finally:
        ctx.check(ctx.lib.dcpse_set_sell_pipeline(ctx.h, *(best[0] if best else (2, 2))), "dcpse_set_sell_pipeline")
    return best, results


def tune_assembly(fn, ppws=(1, 2, 4, 8, 16), reps=2, device=0):
    """fn() runs one assembly; returns (particles_per_wave, ms) of the fastest."""
    ctx = _lib.context(device)
    best = None
    results = {}
    for p in ppws:
        ctx.check(ctx.lib.dcpse_set_assembly_config(ctx.h, 1, p), "dcpse_set_assembly_config")
        ms = _time(fn, reps)
        results[p] = ms
        if best is None or ms < best[1]:
            best = (p, ms)
    ctx.check(ctx.lib.dcpse_set_assembly_config(ctx.h, 1, best[0]), "dcpse_set_assembly_config")
    return best, results
