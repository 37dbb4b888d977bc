"""Launch-shape autotuner for the apply and assembly kernels.

Times each candidate with HIP events on the current stream (the stream the
library launches on) and keeps the fastest in the library context:
  apply:    CSR (lanes per row, nonzeros per lane, load width) x threads per block;
            sliced-ELL groups of 4 slots loaded ahead {2, 4}
  assembly: grid (static: a block per batch; persistent: a counter hands out
            batches), particles per batch, neighbour chunk {32, 64} and
            bucketing of the particles by row length (longest first)
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


SELL_PIPELINES = ((2, 2), (4, 2))  # (groups per batch, stages)


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


#: the default assembly launch: static grid, 18 particles per batch (the
#: kernels factor two or three particles together), chunks of 64 neighbours,
#: index order
ASSEMBLY_DEFAULT = (0, 18, 64, 0)

#: (persistent, particles per batch, chunk, bucket width) tried by tune_assembly
ASSEMBLY_SPACE = tuple((pers, ppb, ch, bw) for pers in (0, 1) for ppb in (6, 12, 18) for ch in (32, 64)
                       for bw in ((0, 16) if pers else (0,)))


def set_assembly_launch(persistent=0, particles_per_batch=18, chunk=64, bucket_width=0, device=0):
    """Launch shape of the assembly kernels (dcpse_set_assembly_launch); the
    results do not depend on it."""
    ctx = _lib.context(device)
    ctx.check(ctx.lib.dcpse_set_assembly_launch(ctx.h, int(persistent), int(particles_per_batch), int(chunk),
                                                int(bucket_width)), "dcpse_set_assembly_launch")


def tune_assembly(fn, space=ASSEMBLY_SPACE, reps=2, device=0):
    """fn() runs one assembly; every launch shape of ``space`` (persistent,
    particles per batch, chunk, bucket width) is timed with HIP events and the
    fastest is kept in the library context. Returns ((shape), ms), results."""
    best = None
    results = {}
    for cfg in space:
        set_assembly_launch(*cfg, device=device)
        ms = _time(fn, reps)
        results[cfg] = ms
        if best is None or ms < best[1]:
            best = (cfg, ms)
    set_assembly_launch(*best[0], device=device)
    return best, results
