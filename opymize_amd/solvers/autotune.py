"""Tile-shape, iterations-per-launch and iterations-per-graph selection for
the fused TV kernels.

`default_tile` is a shape heuristic; `tune` times candidate (tile width,
rows per thread, iterations per launch) triples and graph block lengths on
the actual batch for a few iterations and keeps the fastest.  Results are
cached per process keyed by (C, dtype, size bucket, passes, step rule) and
optionally persisted to a JSON file.
"""
import json
import os
import time

import numpy as np
import torch

TILE_WIDTHS = (32, 64, 128, 256)          # two-pass kernels
ROWS_PER_THREAD = (1, 2, 4, 8)
ITER_ROWS_PER_THREAD = (4, 6, 8, 12, 16)  # single-pass kernel (tile width 64)
MULTI_ROWS_PER_THREAD = (2, 4, 6, 8)         # several iterations per launch
MULTI = (1, 2, 3, 4)                      # iterations per launch (constant / precond / accel steps)
BLOCK_ITERS = (12, 24, 48, 96)

_cache = {}


def default_tile(H, W, C, passes=1):
    """Single pass: 64 columns, 8 rows per thread for C=1 and 4 beyond (register
    budget).  Two passes: full wavefront rows for wide images, narrower tiles
    for thin ones."""
    if passes == 1:
        return 64, (8 if C == 1 else 4)
    tw = 64
    while tw > 8 and tw // 2 >= W:
        tw //= 2
    tw = max(tw, 8)
    return tw, 4


def _key(batch, steps):
    px = sum(h * w for h, w in batch.shapes)
    return (batch.C, str(batch.dtype), int(np.log2(max(px, 1))), batch.B > 1, batch.passes, steps == "adaptive")


def _apply(batch, best):
    batch.set_multi(best.get("multi", 1))
    batch.set_tile(*best["tile"])


def tune(batch, steps="adaptive", step_bound=0.124, iters=24, widths=None, rpts=None, multis=None,
         blocks=BLOCK_ITERS, path=None, verbose=False):
    """Pick the fastest (tw, rpt, iterations per launch) for `batch` and the
    graph block length; leaves the batch set to the winner and returns
    {"tile": ..., "multi": ..., "block_iters": ...}.  Several iterations per
    launch are candidates with single-pass kernels and non-adaptive steps;
    shapes whose LDS region does not fit are skipped."""
    from opymize_amd import _hip
    if widths is None:
        widths = (64,) if batch.passes == 1 else TILE_WIDTHS
    if multis is None:
        multis = MULTI if (batch.passes == 1 and steps != "adaptive") else (1,)
    key = _key(batch, steps)
    if key in _cache:
        best = _cache[key]
        _apply(batch, best)
        return best
    if path and os.path.exists(path):
        with open(path) as fh:
            stored = json.load(fh)
        if str(key) in stored:
            best = stored[str(key)]
            best["tile"] = tuple(best["tile"])
            _apply(batch, best)
            _cache[key] = best
            return best
    x0 = batch.x[0].clone()
    y0 = batch.y[0].clone()
    timings = {}
    for m in multis:
        if rpts is not None:
            cand = rpts
        elif m > 1:
            cand = MULTI_ROWS_PER_THREAD
        else:
            cand = ITER_ROWS_PER_THREAD if batch.passes == 1 else ROWS_PER_THREAD
        for tw in widths:
            for rpt in cand:
                batch.set_multi(m)
                batch.set_tile(tw, rpt)
                batch.set_start(x0, y0)
                n = max(m, iters - iters % m)
                try:
                    batch.solve(steps=steps, step_bound=step_bound, max_iters=m, block_iters=m, use_graph=False,
                                tol=0.0)
                except _hip.HipError:  # region beyond the LDS budget
                    continue
                torch.cuda.synchronize()
                t = time.perf_counter()
                batch.solve(steps=steps, step_bound=step_bound, max_iters=n, block_iters=n, use_graph=False,
                            tol=0.0)
                torch.cuda.synchronize()
                timings[(tw, rpt, m)] = (time.perf_counter() - t) / n
                if verbose:
                    print("tile %3dx%d, %d per launch: %.3f ms/iter" % (tw, rpt, m, 1e3 * timings[(tw, rpt, m)]),
                          flush=True)
    tw, rpt, m = min(timings, key=timings.get)
    tile = (tw, rpt)
    batch.set_multi(m)
    batch.set_tile(*tile)
    btime = {}
    for nb in blocks:
        batch.set_start(x0, y0)
        batch.solve(steps=steps, step_bound=step_bound, max_iters=nb, block_iters=nb, tol=0.0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        batch.solve(steps=steps, step_bound=step_bound, max_iters=4 * nb, block_iters=nb, tol=0.0,
                    keep_state=False)
        torch.cuda.synchronize()
        btime[nb] = (time.perf_counter() - t) / (4 * nb)
    block = min(btime, key=btime.get)
    batch.set_start(x0, y0)
    best = {"tile": tile, "multi": int(m), "block_iters": int(block), "ms_per_iter": 1e3 * timings[(tw, rpt, m)]}
    _cache[key] = best
    if path:
        stored = {}
        if os.path.exists(path):
            with open(path) as fh:
                stored = json.load(fh)
        stored[str(key)] = {"tile": list(tile), "multi": int(m), "block_iters": int(block),
                            "ms_per_iter": best["ms_per_iter"]}
        with open(path, "w") as fh:
            json.dump(stored, fh, indent=1)
    return best
