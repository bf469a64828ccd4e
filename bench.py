"""Throughput of the PDHG hot path on one MI355X.

Metric: solves/sec to a relative primal-dual gap of 1e-5, certified in FP64,
for the batched (and, when a trained warm-start network is available,
warm-started) fused PDHG on the learned-warm-start workload: TV-L2 denoising
of 1024x1024 images from the rectangle generator, lambda log-uniform in
[1, 20], 256 instances per batch.  A step = solving one resident batch from
its start point until every instance is certified or the iteration cap.

Printed next to it: the plain single-instance GPU PDHG (the drop-in generic
kernel chain of PDHG.solve, cold start, same step rule) on a few instances of
the same batch; the roofline of the dominant kernel (HIP events on its
stream); and the CPU oracle timed on the host (bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_batch(args):
    import torch
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    specs = synth.batch_square(args.batch, side=args.size, seed=args.seed, lam_lo=1.0, lam_hi=20.0, data="l2")
    t = time.time()
    imgs = synth.images(specs)
    log("generated %d images of %dx%d in %.1fs" % (len(imgs), args.size, args.size, time.time() - t))
    bat = TVBatch(imgs, [s["lam"] for s in specs], "l2", precision="float", tile=tuple(args.tile))
    del imgs
    torch.cuda.synchronize()
    return bat, specs


def time_dominant_kernel(bat, reps=10):
    """Average duration of the fused dual pass over the full batch (all
    instances running), timed with HIP events on the launch stream."""
    import ctypes
    import torch
    from opymize_amd import _hip
    steps = "adaptive"
    cfg = bat._cfg(steps)
    bat._reset_state(steps, 0.124, 1.0, 1e-5, 1e-5, True)
    P, s, sfx = _hip.ptr, _hip.stream(), bat._sfx()
    stream = torch.cuda.current_stream()

    def launch(name):
        if name == "dual":
            bat._call("opy_tv_dual" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.insts),
                      P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact), ctypes.byref(cfg), 1,
                      P(bat.part_res), bat.grid, s)
        else:
            bat._call("opy_tv_primal" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.f), P(bat.insts),
                      P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact), ctypes.byref(cfg), bat.grid, s)

    out = {}
    for name in ("dual", "primal"):
        launch(name)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch(name)
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    return out


def plain_single_instance(bat, specs, n, max_iters):
    """Drop-in generic PDHG (per-operator kernels, graph-captured) on n instances."""
    import torch
    from opymize_amd.functionals import SSD, L1Norms
    from opymize_amd.linear import GradientOp
    from opymize_amd.solvers import PDHG
    f_all = bat.unpack(bat.f, "x")
    solved, iters, t_tot = 0, [], 0.0
    for b in range(n):
        H, W = bat.shapes[b]
        f = f_all[b].reshape(-1, 1)
        G = SSD(f, vol=np.array([specs[b]["lam"]]))
        F = L1Norms(H * W, (1, 2), 1.0)
        A = GradientOp((H, W), 1, scheme="forward")
        s = PDHG(G, F, A)
        torch.cuda.synchronize()
        t = time.perf_counter()
        info = s.solve(steps="adaptive", precision="float", step_bound=0.124, term_maxiter=max_iters,
                       granularity=50, term_pd_gap=1e-5, fused=False)
        torch.cuda.synchronize()
        t_tot += time.perf_counter() - t
        ok = abs(info["relgap"]) < 1e-5
        solved += ok
        iters.append(info["iter"])
    return {"value": solved / t_tot, "instances": n, "certified": solved, "iters": iters, "seconds": t_tot}


def cpu_baseline(bat, specs, budget_s):
    """Oracle (float64 numpy/scipy, 1 thread) on instance 0: time per
    iteration over a bounded sample, scaled by the iterations the GPU needed
    for that instance."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import pdhg as opd
    f = bat.unpack(bat.f, "x")[0].reshape(-1, 1)
    H, W = bat.shapes[0]
    prob = opd.tv_problem(f, (H, W), specs[0]["lam"])
    n_it = 5
    t = time.perf_counter()
    opd.solve(prob, steps="adaptive", step_bound=0.124, term_maxiter=n_it, granularity=10 ** 9)
    dt = time.perf_counter() - t
    while dt < budget_s and n_it < 400:
        n_it = max(n_it * 2, int(n_it * budget_s / max(dt, 1e-3)))
        n_it = min(n_it, 400)
        t = time.perf_counter()
        opd.solve(prob, steps="adaptive", step_bound=0.124, term_maxiter=n_it, granularity=10 ** 9)
        dt = time.perf_counter() - t
    return dt / n_it, n_it, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--max-iters", type=int, default=4000)
    ap.add_argument("--block-iters", type=int, default=50)
    ap.add_argument("--tile", type=int, nargs=2, default=[64, 4])
    ap.add_argument("--plain", type=int, default=2, help="instances for the plain single-instance baseline")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (instances would shard across "
              "GPUs by bin-packing, see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from opymize_amd import _hip
    _hip.require_device()
    _hip.load()
    bat, specs = build_batch(args)
    npx = sum(h * w for h, w in bat.shapes)

    def one_step():
        bat.set_start(None, None)
        return bat.solve(steps="adaptive", step_bound=0.124, max_iters=args.max_iters,
                         block_iters=args.block_iters, tol=1e-5, certify=True)

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    certified = 0
    iters = []
    for _ in range(args.steps):
        res = one_step()
        certified += sum(r["status"] == 1 for r in res)
        iters += [r["iter"] for r in res]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = certified / elapsed
    iters = np.array(iters)
    log("batched: %d/%d certified in %.2fs, iterations min/median/max %d/%d/%d"
        % (certified, args.steps * args.batch, elapsed, iters.min(), np.median(iters), iters.max()))

    kt = time_dominant_kernel(bat)
    dual_bytes = 24.0 * npx  # reads x, x+, y (4+4+8 B), writes y+ (8 B) per pixel
    achieved = dual_bytes / (kt["dual"] * 1e-3) / 1e9
    log("dual pass %.3f ms (%.0f GB/s), primal pass %.3f ms (%.0f GB/s)"
        % (kt["dual"], achieved, kt["primal"], 20.0 * npx / (kt["primal"] * 1e-3) / 1e9))

    plain = plain_single_instance(bat, specs, args.plain, args.max_iters) if args.plain > 0 else None
    if plain:
        log("plain single-instance: %d/%d certified, %.3f solves/s, iterations %s"
            % (plain["certified"], plain["instances"], plain["value"], plain["iters"]))

    cpu = None
    if not args.no_cpu:
        per_it, n_it, dt = cpu_baseline(bat, specs, args.cpu_budget)
        it0 = int(iters[0]) if len(iters) else args.max_iters
        cpu = {"value": 1.0 / (per_it * it0), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": "oracle (float64 numpy/scipy, adaptive steps) on instance 0 (1024x1024): %d iterations "
                         "in %.1fs, scaled to the %d iterations the GPU needed for that instance"
                         % (n_it, dt, it0)}
        log("cpu oracle: %.3f s/iteration" % per_it)

    line = {
        "metric": "solves/sec to rel PD-gap 1e-5: batched+warm-start vs plain PDHG; %HBM peak",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "TV-L2 denoising, %d instances of %dx%d per batch, rectangle generator, lambda "
                               "log-uniform [1,20], adaptive steps, rel gap 1e-5 FP64-certified, cap %d iterations"
                               % (args.batch, args.size, args.size, args.max_iters),
                   "warm_start": "none", "tile": list(args.tile), "block_iters": args.block_iters,
                   "certified": int(certified), "instances": int(args.steps * args.batch),
                   "iterations_median": float(np.median(iters)),
                   "plain_single_instance_solves_per_sec": plain["value"] if plain else None,
                   "ratio_vs_plain": (value / plain["value"]) if plain and plain["value"] > 0 else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "tv_dual_kernel<float,1,adaptive> (24 B/pixel algorithmic)",
                     "avg_ms": kt["dual"], "primal_avg_ms": kt["primal"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
