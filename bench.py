"""Throughput of the PDHG hot path on one MI355X.

Metric: solves/sec to a relative primal-dual gap of 1e-5, certified in FP64,
for the batched fused PDHG on the learned-warm-start workload: TV-L2
denoising of 1024x1024 images from the rectangle generator, lambda
log-uniform in [1, 20], 256 instances per batch.  A step = the trained
network's prediction of (x, y) for the resident batch, then the batched
solver from that start until every instance is certified or the iteration
cap, with a cold restart for any instance whose gap does not improve at the
first check (--cold: from x = y = 0 instead).  The batched solver runs the
accelerated steps of Chambolle & Pock (2011, Algorithm 2) on these strongly
convex instances, restarting an instance's step schedule when its certified
gap has fallen enough or stopped falling (--no-restart: never; --step-rule
adaptive: the reference's rule).  After the
timed steps the other start (cold, or warm with --cold) is solved once more
for comparison.

Printed next to it:
  * the plain single-instance GPU PDHG (the drop-in generic kernel chain of
    PDHG.solve, cold start, the reference's adaptive steps, same cap) on
    instances spread over the lambda range of the same batch;
  * the roofline of the dominant kernel (by default the kernel that runs K
    iterations per launch on LDS-staged tiles; the single-pass iteration
    kernel with --multi 1 or adaptive steps; the dual pass with --passes 2):
    algorithmic bytes of a full-batch launch over the average duration of
    eager full-batch launches, timed with HIP events on the launch stream;
  * the CPU oracle (float64 numpy/scipy, one thread) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cold]
       python bench.py --roofline-only   (only the kernel-timing phase; for counter runs)
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
ITER_BYTES_PER_PX = 28.0    # single pass, C=1 FP32: reads x_k, f (4+4 B), y_k (8 B); writes x_k+1 (4 B), y_k+1 (8 B)
MULTI_BYTES_PER_PX = 28.0   # K iterations per launch: reads x_k, f, y_k and writes x_k+K, y_k+K once
DUAL_BYTES_PER_PX = 24.0    # two passes: the dual pass reads x_k, x_k+1 (4+4 B), y_k (8 B); writes y_k+1 (8 B)
PRIMAL_BYTES_PER_PX = 20.0  # two passes: the primal pass reads x_k, f (4+4 B), y_k (8 B); writes x_k+1 (4 B)


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
    bat = TVBatch(imgs, [s["lam"] for s in specs], "l2", precision="float", tile=tuple(args.tile),
                  passes=args.passes,
                  multi=args.multi if (args.step_rule != "adaptive" and args.passes == 1) else 1)
    del imgs
    torch.cuda.synchronize()
    return bat, specs


def roofline_phase(bat, reps, rule):
    """Eager full-batch launches of the iteration kernel(s) with the bench's
    step rule (every instance running), timed with HIP events on the stream
    they are launched on; one launch of the library's copy kernel over a
    buffer of known size for counter scaling."""
    import ctypes
    import torch
    from opymize_amd import _hip
    cfg = bat._cfg(rule)
    adapt = int(rule == "adaptive")
    bat._reset_state(rule, 0.124, 1.0, 1e-5, 1e-5, True)
    bat.refresh_grid()  # every instance running: one block per tile of the batch
    P, sfx = _hip.ptr, bat._sfx()
    stream = torch.cuda.current_stream()
    s = _hip.stream()

    def launch(name):
        if name == "multi":
            bat._call("opy_tv_multi" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.f),
                      P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact),
                      ctypes.byref(cfg), bat.multi, bat.grid, s)
        elif name == "iter":
            bat._call("opy_tv_iter" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.f),
                      P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact),
                      ctypes.byref(cfg), adapt, P(bat.part_res), bat.grid, s)
        elif name == "dual":
            bat._call("opy_tv_dual" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.insts),
                      P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact), ctypes.byref(cfg), adapt,
                      P(bat.part_res), bat.grid, s)
        else:
            bat._call("opy_tv_primal" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.f), P(bat.insts),
                      P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact), ctypes.byref(cfg), bat.grid, s)

    out = {"grid": bat.grid}
    if bat.passes == 2:
        names = ("dual", "primal")
    else:
        names = ("multi",) if (bat.multi > 1 and rule != "adaptive") else ("iter",)
    for name in names:
        launch(name)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch(name)
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    # copy of a known size (2^28 floats = 1 GiB read + 1 GiB written): the counter scale
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    _hip.call("copy", a.dtype, P(a), P(b), a.numel(), s)
    torch.cuda.synchronize()
    del a, b
    return out


def measured_traffic(npx, tile, passes, rule, multi):
    """HBM bytes per full-batch launch of the dominant kernel from the
    committed counter runs (FETCH_SIZE and WRITE_SIZE, each scaled by what it
    reads for a copy of known size), when they were taken on this workload;
    else None."""
    if passes == 2:
        name = "tv_dual_kernel"
    elif multi > 1:
        name = "tv_multi_kernel_k%d" % multi
    else:
        name = "tv_iter_kernel"
    name += "_accel" if rule == "accel" else ""
    path = os.path.join(ROOT, "profiles", "%s_tvl2_256x1024_counters.json" % name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        m = json.load(fh)
    if m["workload"]["pixels"] != npx or list(m["workload"]["tile"]) != list(tile):
        return None
    kib = 1024.0
    fetch = m["fetch_kib"] * (m["copy_known_bytes_each_way"] / (m["copy_fetch_kib"] * kib))
    write = m["write_kib"] * (m["copy_known_bytes_each_way"] / (m["copy_write_kib"] * kib))
    return (fetch + write) * kib


def plain_single_instance(bat, specs, n, max_iters):
    """Drop-in generic PDHG (per-operator kernels, graph-captured), cold start,
    on n instances spread evenly over the lambda order of the batch."""
    import torch
    from opymize_amd.functionals import SSD, L1Norms
    from opymize_amd.linear import GradientOp
    from opymize_amd.solvers import PDHG
    f_all = bat.unpack(bat.f, "x")
    order = np.argsort([s["lam"] for s in specs])
    pick = [int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, n)]
    solved, iters, t_tot = 0, [], 0.0
    for b in pick:
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
        solved += abs(info["relgap"]) < 1e-5
        iters.append(int(info["iter"]))
    return {"value": solved / t_tot, "instances": pick, "certified": int(solved), "iters": iters,
            "seconds": t_tot, "lams": [round(specs[b]["lam"], 3) for b in pick]}


def cpu_baseline(bat, specs, budget_s, gpu_iters, rule):
    """Oracle (float64 numpy/scipy, one thread, same step rule) on instance 0:
    seconds per iteration over a bounded sample, scaled by the iterations the
    GPU needed for that instance."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import pdhg as opd
    f = bat.unpack(bat.f, "x")[0].reshape(-1, 1)
    H, W = bat.shapes[0]
    prob = opd.tv_problem(f, (H, W), specs[0]["lam"])
    n_it = 5
    t = time.perf_counter()
    opd.solve(prob, steps=rule, step_bound=0.124, term_maxiter=n_it, granularity=10 ** 9)
    dt = time.perf_counter() - t
    while dt < budget_s and n_it < 400:
        n_it = min(400, max(n_it * 2, int(n_it * budget_s / max(dt, 1e-3))))
        t = time.perf_counter()
        opd.solve(prob, steps=rule, step_bound=0.124, term_maxiter=n_it, granularity=10 ** 9)
        dt = time.perf_counter() - t
    per_it = dt / n_it
    return {"value": 1.0 / (per_it * gpu_iters), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle (float64 numpy/scipy, %s steps, 1 thread) on instance 0 (%dx%d): %d iterations "
                      "in %.1fs, scaled to the %d iterations the GPU needed for that instance"
                      % (rule, H, W, n_it, dt, gpu_iters)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--max-iters", type=int, default=10000)
    ap.add_argument("--block-iters", type=int, default=48)
    ap.add_argument("--tile", type=int, nargs=2, default=[64, 6])
    ap.add_argument("--multi", type=int, default=4, choices=[1, 2, 3, 4],
                    help="iterations per launch (accelerated steps; adaptive steps take 1)")
    ap.add_argument("--passes", type=int, default=1, choices=[1, 2])
    ap.add_argument("--cold", action="store_true", help="start from x = y = 0 instead of the network's prediction")
    ap.add_argument("--step-factor", type=float, default=4.0,
                    help="initial tau/sigma split: tau = sqrt(bound * factor), sigma = sqrt(bound / factor)")
    ap.add_argument("--no-compare", action="store_true", help="skip the solve from the other start")
    ap.add_argument("--no-restart", action="store_true",
                    help="accelerated steps without restarts of the step schedule")
    ap.add_argument("--step-rule", default="accel", choices=["accel", "adaptive"],
                    help="step rule of the batched solver")
    ap.add_argument("--plain", type=int, default=6, help="instances for the plain single-instance baseline")
    ap.add_argument("--reps", type=int, default=20, help="launches per pass in the roofline phase")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: opymize_amd has no multi-GPU path (a batch would shard across GPUs by "
              "instance, each GPU solving its own share); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from opymize_amd import _hip
    _hip.require_device()
    _hip.load()
    bat, specs = build_batch(args)
    npx = sum(h * w for h, w in bat.shapes)
    lams = [s["lam"] for s in specs]

    if args.roofline_only:
        kt = roofline_phase(bat, args.reps, args.step_rule)
        log("roofline phase: %s" % ", ".join("%s %s" % kv for kv in kt.items()))
        return

    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    net = WarmStart(ConvNet.load(os.path.join(ROOT, "opymize_amd", "warmstart", "weights_tv_l2.npz")))

    def one_step(warm):
        if warm:
            net.predict_into(bat, lams)
        else:
            bat.set_start(None, None)
        return bat.solve(steps=args.step_rule, step_bound=0.124, step_factor=args.step_factor,
                         max_iters=args.max_iters, block_iters=args.block_iters, tol=1e-5, certify=True,
                         fallback=warm, restart=not args.no_restart)

    warm = not args.cold

    for _ in range(args.warmup):
        one_step(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    certified = 0
    iters = []
    fallbacks = 0
    nrestarts = []
    for _ in range(args.steps):
        res = one_step(warm)
        fallbacks += len(bat.fallbacks)
        certified += sum(r["status"] == 1 for r in res)
        iters += [r["iter"] for r in res]
        nrestarts += [r["restarts"] for r in res]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = certified / elapsed
    iters = np.array(iters)
    log("batched (%s start): %d/%d certified in %.2fs, iterations min/median/max %d/%d/%d"
        % ("warm" if warm else "cold", certified, args.steps * args.batch, elapsed, iters.min(), np.median(iters),
           iters.max()))
    other = None
    if not args.no_compare:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = one_step(not warm)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        c = int(sum(r["status"] == 1 for r in res))
        other = {"start": "cold" if warm else "warm", "certified": c, "seconds": dt, "solves_per_sec": c / dt,
                 "iterations_median": float(np.median([r["iter"] for r in res]))}
        log("batched (%s start, for comparison): %d/%d certified in %.2fs" % (other["start"], c, args.batch, dt))

    kt = roofline_phase(bat, args.reps, args.step_rule)
    variant = "adaptive" if args.step_rule == "adaptive" else "accelerated steps"
    k_launch = 1
    if bat.passes == 2:
        kname, kms, kbytes = "dual", kt["dual"], DUAL_BYTES_PER_PX
        kdesc = "tv_dual_kernel<float,1,%d>, %s (second of two passes)" % (args.tile[1], variant)
    elif "multi" in kt:
        k_launch = bat.multi
        kname, kms, kbytes = "multi", kt["multi"], MULTI_BYTES_PER_PX
        kdesc = "tv_multi_kernel<float,1,%d,%d>, %s (%d iterations per launch on LDS-staged tiles)" % (
            args.tile[1], k_launch, variant, k_launch)
    else:
        kname, kms, kbytes = "iter", kt["iter"], ITER_BYTES_PER_PX
        kdesc = "tv_iter_kernel<float,1,%d>, %s (one pass per iteration)" % (args.tile[1], variant)
    achieved = kbytes * npx / (kms * 1e-3) / 1e9
    log("roofline phase: %s -- %s %.0f GB/s algorithmic" % (
        ", ".join("%s %s" % kv for kv in kt.items()), kname, achieved))

    plain = plain_single_instance(bat, specs, args.plain, args.max_iters) if args.plain > 0 else None
    if plain:
        log("plain single-instance: %d/%d certified, %.3f solves/s, iterations %s, lambdas %s"
            % (plain["certified"], len(plain["instances"]), plain["value"], plain["iters"], plain["lams"]))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bat, specs, args.cpu_budget, int(iters[0]), args.step_rule)
        log("cpu oracle: %s" % cpu["sample"])

    from opymize_amd.solvers import tv as tvmod
    line = {
        "metric": "solves/sec to rel PD-gap 1e-5: batched+warm-start vs plain PDHG; %HBM peak",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "TV-L2 denoising, %d instances of %dx%d per batch, rectangle generator, lambda "
                               "log-uniform [1,20], rel gap 1e-5 FP64-certified, cap %d iterations"
                               % (args.batch, args.size, args.size, args.max_iters),
                   "step_rule": {"accel": "accelerated (Chambolle-Pock Alg. 2, gamma = lambda)",
                                 "adaptive": "adaptive (reference rule)"}[args.step_rule],
                   "plain_step_rule": "adaptive (reference rule)",
                   "warm_start": ("cnn (6-layer 3x3 conv, 32 channels, bf16 MFMA), cold restart where the gap "
                                  "does not improve (%d instances)" % fallbacks) if warm else "none (cold start)",
                   "step_factor": args.step_factor,
                   "restarts": None if (args.no_restart or args.step_rule != "accel") else
                   {"beta": tvmod.RESTART_DEFAULT[0], "min_iters": tvmod.RESTART_DEFAULT[1],
                    "median_per_instance": float(np.median(nrestarts))},
                   "other_start": other,
                   "tile": list(args.tile), "block_iters": args.block_iters, "iterations_per_launch": k_launch,
                   "certified": int(certified), "instances": int(args.steps * args.batch),
                   "iterations_median": float(np.median(iters)), "iterations_max": int(iters.max()),
                   "plain_single_instance_solves_per_sec": plain["value"] if plain else None,
                   "plain_instances": len(plain["instances"]) if plain else 0,
                   "ratio_vs_plain": (value / plain["value"]) if plain and plain["value"] > 0 else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(npx, args.tile, args.passes, args.step_rule, k_launch),
                     "kernel": "%s: %d B/pixel algorithmic x %d pixels per launch" % (kdesc, kbytes, npx),
                     "avg_ms": kms, "ms_per_iteration": kms / k_launch,
                     "launch_ms": {k: v for k, v in kt.items() if k != "grid"},
                     "blocks_per_launch": kt["grid"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
