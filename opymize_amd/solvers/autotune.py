"""Tile-shape and iterations-per-graph selection for the fused TV kernels.

`default_tile` is a shape heuristic; `tune` times candidate (tile width,
rows per thread) pairs and graph block lengths on the actual batch for a few
iterations and keeps the fastest.  Results are cached per process keyed by
(C, dtype, size bucket) and optionally persisted to a JSON file.
"""
import json
import os
import time

import numpy as np
import torch

TILE_WIDTHS = (32, 64, 128, 256)          # two-pass kernels
ROWS_PER_THREAD = (1, 2, 4, 8)
ITER_ROWS_PER_THREAD = (4, 6, 8, 12, 16)  # single-pass kernel (tile width 64)
BLOCK_ITERS = (10, 20, 50, 100)

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


def _key(batch):
    px = sum(h * w for h, w in batch.shapes)
    return (batch.C, str(batch.dtype), int(np.log2(max(px, 1))), batch.B > 1, batch.passes)


def tune(batch, steps="adaptive", step_bound=0.124, iters=20, widths=None, rpts=None,
         blocks=BLOCK_ITERS, path=None, verbose=False):
    """Pick the fastest (tw, rpt) for `batch` and the graph block length;
    leaves the batch set to the winner and returns {"tile": ..., "block_iters": ...}."""
    if widths is None:
        widths = (64,) if batch.passes == 1 else TILE_WIDTHS
    if rpts is None:
        rpts = ITER_ROWS_PER_THREAD if batch.passes == 1 else ROWS_PER_THREAD
    key = _key(batch)
    if key in _cache:
        best = _cache[key]
        batch.set_tile(*best["tile"])
        return best
    if path and os.path.exists(path):
        with open(path) as fh:
            stored = json.load(fh)
        if str(key) in stored:
            best = stored[str(key)]
            best["tile"] = tuple(best["tile"])
            batch.set_tile(*best["tile"])
            _cache[key] = best
            return best
    x0 = batch.x[0].clone()
    y0 = batch.y[0].clone()
    timings = {}
    for tw in widths:
        for rpt in rpts:
            batch.set_tile(tw, rpt)
            batch.set_start(x0, y0)
            batch.solve(steps=steps, step_bound=step_bound, max_iters=2, block_iters=2, use_graph=False,
                        tol=0.0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            batch.solve(steps=steps, step_bound=step_bound, max_iters=iters, block_iters=iters, use_graph=False,
                        tol=0.0)
            torch.cuda.synchronize()
            timings[(tw, rpt)] = (time.perf_counter() - t) / iters
            if verbose:
                print("tile %3dx%d: %.3f ms/iter" % (tw, rpt, 1e3 * timings[(tw, rpt)]), flush=True)
    tile = min(timings, key=timings.get)
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
    best = {"tile": tile, "block_iters": int(block), "ms_per_iter": 1e3 * timings[tile]}
    _cache[key] = best
    if path:
        stored = {}
        if os.path.exists(path):
            with open(path) as fh:
                stored = json.load(fh)
        stored[str(key)] = {"tile": list(tile), "block_iters": int(block), "ms_per_iter": best["ms_per_iter"]}
        with open(path, "w") as fh:
            json.dump(stored, fh, indent=1)
    return best
