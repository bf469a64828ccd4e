"""Timing of the generic PDHG path (one kernel per operator, no fused TV
kernels) on a TV-L2 image, for profiles: a warm-up solve, then a timed solve
of `iters` iterations per step rule, and the time per iteration.

Usage: python -m opymize_amd.tools.pdhg_bench [--side S] [--iters K] [--precision float|double]
"""
import argparse
import time


def main(argv=None):
    import numpy as np
    import torch
    from opymize_amd.functionals import SSD, L1Norms
    from opymize_amd.linear import GradientOp
    from opymize_amd.solvers import PDHG
    from opymize_amd.tools import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--precision", default="float", choices=["float", "double"])
    a = ap.parse_args(argv)
    H = W = a.side
    f = synth.rects(H, W, 1, seed=3).reshape(-1, 1)
    G = SSD(f, vol=np.array([4.0]))
    F = L1Norms(H * W, (1, 2), lbd=1.0)
    A = GradientOp((H, W), 1, scheme="forward")
    for steps in ("adaptive", "const"):
        kw = dict(steps=steps, precision=a.precision, step_bound=0.124, fused=False)
        PDHG(G, F, A).solve(term_maxiter=50, granularity=50, **kw)
        torch.cuda.synchronize()
        t = time.time()
        PDHG(G, F, A).solve(term_maxiter=a.iters, granularity=a.iters, **kw)
        torch.cuda.synchronize()
        print("%s steps, %dx%d TV-L2 (%s): %.1f us per iteration"
              % (steps, H, W, a.precision, (time.time() - t) / a.iters * 1e6), flush=True)


if __name__ == "__main__":
    main()
