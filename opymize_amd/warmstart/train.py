"""Training of the warm-start network on pairs the solver generates.

1. Solve a batch of TV-L2 instances (rectangle generator, lambda log-uniform
   in [1, 20]) with the batched fused PDHG to a tight relative gap.
2. Cut random crops of (f, lambda) -> (x* - f, y*) from the solutions.
3. Train the network with Adam on the matrix-core conv kernels.
4. Evaluate: solve a held-out batch cold and warm-started (with cold-start
   fallback) and report iterations, time and the gap of the start points.

Usage: python -m opymize_amd.warmstart.train --out weights.npz [options]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from opymize_amd.solvers.tv import TVBatch
from opymize_amd.tools import synth
from opymize_amd.warmstart.cnn import CH, ConvNet, WarmStart


def log(*a):
    print(*a, flush=True)


def solve_batch(specs, tol, max_iters, warm=None):
    imgs = synth.images(specs)
    lams = [s["lam"] for s in specs]
    bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, 8))
    t_inf = 0.0
    if warm is not None:
        torch.cuda.synchronize()
        t = time.perf_counter()
        warm.predict_into(bat, lams)
        torch.cuda.synchronize()
        t_inf = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = bat.solve(steps="adaptive", step_bound=0.124, tol=tol, max_iters=max_iters, block_iters=50,
                    fallback=warm is not None)
    torch.cuda.synchronize()
    return bat, res, time.perf_counter() - t, t_inf, imgs, lams


def make_crops(bat, imgs, lams, crop, per_image, seed):
    """Device tensors: inputs (N, crop, crop, 2) and targets (N, crop, crop, 3)."""
    rng = np.random.default_rng(seed)
    ins, tgs = [], []
    for b, (H, W) in enumerate(bat.shapes):
        off = bat.offs[b]
        f = bat.f[off:off + H * W].view(H, W)
        x = bat.x[0][off:off + H * W].view(H, W)
        y = bat.y[0][2 * off:2 * off + 2 * H * W].view(H, W, 2)
        for _ in range(per_image):
            r0, c0 = int(rng.integers(0, H - crop + 1)), int(rng.integers(0, W - crop + 1))
            fc = f[r0:r0 + crop, c0:c0 + crop]
            i = torch.empty((crop, crop, 2), dtype=torch.float32, device="cuda")
            i[..., 0] = fc - 0.5
            i[..., 1] = float(np.log(lams[b]) / 3.0)
            t = torch.empty((crop, crop, 3), dtype=torch.float32, device="cuda")
            t[..., 0] = x[r0:r0 + crop, c0:c0 + crop] - fc
            t[..., 1:3] = y[r0:r0 + crop, c0:c0 + crop]
            ins.append(i)
            tgs.append(t)
    return torch.stack(ins), torch.stack(tgs)


def train(net, inputs, targets, epochs, batch, lr, wts, seed=0):
    n, S = inputs.shape[0], inputs.shape[1]
    rng = np.random.default_rng(seed)
    xb = torch.zeros((batch, S, S, CH), dtype=torch.float32, device="cuda")
    tb = torch.zeros((batch, S, S, CH), dtype=torch.float32, device="cuda")
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot, cnt = 0.0, 0
        lr_e = lr * (0.5 ** (ep // max(1, epochs // 3)))
        for k in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(perm[k:k + batch]).cuda()
            xb[..., :2] = inputs[idx]
            tb[..., :3] = targets[idx]
            part = net.train_step(xb, tb, lr=lr_e, wts=wts)
            tot += part.sum().item()
            cnt += 1
        hist.append(tot / max(cnt, 1))
        log("epoch %d: loss %.5f (lr %.1e)" % (ep, hist[-1], lr_e))
    return hist


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--images", type=int, default=128)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=8000)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--crops-per-image", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--wy", type=float, default=0.25)
    ap.add_argument("--eval-images", type=int, default=64)
    ap.add_argument("--eval-max-iters", type=int, default=6000)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    report = {"args": vars(a)}

    t = time.perf_counter()
    specs = synth.batch_square(a.images, side=a.side, seed=11)
    bat, res, t_solve, _, imgs, lams = solve_batch(specs, a.tol, a.max_iters)
    its = [r["iter"] for r in res]
    report["data"] = {"instances": a.images, "solve_s": t_solve, "certified_1e-6": int(sum(r["status"] == 1 for r in res)),
                      "iterations_median": float(np.median(its)),
                      "relgap_max": float(max(abs(r["relgap_cert"]) for r in res))}
    log("training data: %s" % report["data"])
    inputs, targets = make_crops(bat, imgs, lams, a.crop, a.crops_per_image, seed=1)
    del bat
    torch.cuda.empty_cache()
    log("%d crops of %dx%d" % (inputs.shape[0], a.crop, a.crop))

    net = ConvNet(layers=6, n_in=2, n_out=3, seed=0)
    t = time.perf_counter()
    hist = train(net, inputs, targets, a.epochs, a.batch, a.lr, (1.0, a.wy, a.wy))
    torch.cuda.synchronize()
    report["train"] = {"seconds": time.perf_counter() - t, "loss": hist}
    net.save(a.out)
    del inputs, targets
    torch.cuda.empty_cache()

    # held-out evaluation: cold vs warm
    especs = synth.batch_square(a.eval_images, side=a.side, seed=12)
    _, rc, tc, _, _, _ = solve_batch(especs, 1e-5, a.eval_max_iters)
    bw, rw, tw_, tinf, _, _ = solve_batch(especs, 1e-5, a.eval_max_iters, warm=WarmStart(net))
    ev = {"cold": {"seconds": tc, "certified": int(sum(r["status"] == 1 for r in rc)),
                   "iterations": [r["iter"] for r in rc]},
          "warm": {"seconds": tw_ + tinf, "inference_s": tinf, "certified": int(sum(r["status"] == 1 for r in rw)),
                   "iterations": [r["iter"] for r in rw], "fallbacks": [int(b) for b in bw.fallbacks]},
          "lams": [s["lam"] for s in especs]}
    report["eval"] = ev
    log("eval cold: %.2fs, %d certified, median iters %d" % (tc, ev["cold"]["certified"],
                                                             np.median(ev["cold"]["iterations"])))
    log("eval warm: %.2fs (+%.2fs inference), %d certified, median iters %d, fallbacks %d"
        % (tw_, tinf, ev["warm"]["certified"], np.median(ev["warm"]["iterations"]), len(bw.fallbacks)))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    sys.exit(main())
