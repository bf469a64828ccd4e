"""Timing of the learned warm start's inference (the packed path of
opymize_amd.warmstart.WarmStart: first layer from the packed images, inner
3x3 conv layers on BF16 activations, last layer into the packed x, y starts)
on a batch of TV-L2 instances, for profiles: one warm-up prediction, then
`reps` timed ones, and the time per batch.

Usage: python -m opymize_amd.tools.warmstart_bench [--batch B] [--side S] [--reps R]
"""
import argparse
import os
import time


def main(argv=None):
    import torch
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    specs = synth.batch_square(a.batch, side=a.side, seed=5, lam_lo=1.0, lam_hi=20.0, data="l2")
    lams = [s["lam"] for s in specs]
    bat = TVBatch(synth.images(specs), lams, "l2", precision="float", tile=(64, 6))
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ws = WarmStart(ConvNet.load(os.path.join(here, "warmstart", "weights_tv_l2.npz")))
    ws.predict_into(bat, lams)
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(a.reps):
        ws.predict_into(bat, lams)
    torch.cuda.synchronize()
    dt = (time.time() - t) / a.reps
    print("warm-start inference, %d x %dx%d: %.2f ms per batch (%.2f us per instance)"
          % (a.batch, a.side, a.side, dt * 1e3, dt / a.batch * 1e6), flush=True)


if __name__ == "__main__":
    main()
