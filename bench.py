"""Throughput of the PDHG hot path on one MI355X.

Metric: solves/sec to a relative primal-dual gap of 1e-5, certified in FP64,
for the batched fused PDHG on the learned-warm-start workload: TV-L2
denoising of 1024x1024 images from the rectangle generator, lambda
log-uniform in [1, 20], 256 instances per batch.  A step = the batched
solver from the cold start x = y = 0 until every instance is certified or
the iteration cap (--warm: from the trained network's prediction of (x, y)
for the resident batch, with a cold restart for any instance whose gap does
not improve at the first check).  The cold start is timed by default because
the warm start saves no iterations at 1e-5 under the restarted schedule
below and costs its inference; the warm start is solved once more after the
timed steps and reported beside it (config.other_start).  The batched
solver runs the accelerated steps of Chambolle & Pock (2011, Algorithm 2) on
these strongly convex instances, restarting an instance's step schedule when
its certified gap has fallen enough or stopped falling (--no-restart: never;
--step-rule adaptive: the reference's rule).

--workload lifting instead solves a batch of 512x512, L = 16 multi-label
lifting instances (opymize_amd.lifting) to 1e-5 with the fused batched
kernels (opymize_amd.solvers.lift) and preconditioned steps, against the
plain single-instance drop-in path (the generic kernel chain).

Printed next to it:
  * the plain single-instance GPU PDHG (the drop-in generic kernel chain of
    PDHG.solve, cold start, the reference's adaptive steps, same cap) on
    instances spread over the lambda range of the same batch, and the same
    instances one at a time through the batched solver's own kernels and
    step rule (so that ratio isolates the gain of batching);
  * the roofline of the dominant kernel (by default the streaming kernel
    that runs K = 8 iterations per launch, one wavefront per strip of 48
    columns by 128 rows, the iterations pipelined down the rows; --multi-kernel
    tile: the LDS-tile kernel, K <= 4; the single-pass iteration kernel with
    --multi 1 or adaptive steps; the dual pass with --passes 2): algorithmic
    bytes of a full-batch launch over the average duration of eager
    full-batch launches, timed with HIP events on the launch stream; with
    K > 4 also the same kernel at K = 4 (roofline_k4), where HBM rather than
    instruction issue bounds it;
  * the CPU oracle (float64 numpy/scipy, one thread) on a bounded sample,
    and a pool of processes running it on as many instances at once.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--warm] [--workload tv|lifting]
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
                  multi=args.multi if (args.step_rule != "adaptive" and args.passes == 1) else 1,
                  multi_kernel=getattr(args, "multi_kernel", "auto"), stream_rows=getattr(args, "stream_rows", 128))
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
        if name == "multi" and bat._use_stream(cfg):
            bat._call("opy_tv_stream" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.f),
                      P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.unit_off), P(bat.nact),
                      ctypes.byref(cfg), bat.multi, bat.stream_rows, bat.sgrid, s)
        elif name == "multi":
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

    out = {"grid": bat.sgrid if (bat.multi > 1 and bat._use_stream(cfg)) else bat.grid}
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
    if names == ("multi",) and bat._use_stream(cfg) and bat.multi != 4:
        # the same kernel at 4 iterations per launch, where HBM rather than
        # instruction issue bounds it
        k0 = bat.multi
        bat.set_multi(4)
        bat.refresh_grid()
        launch("multi")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch("multi")
        e1.record(stream)
        torch.cuda.synchronize()
        out["multi_k4"] = e0.elapsed_time(e1) / reps
        out["grid_k4"] = bat.sgrid
        bat.set_multi(k0)
        bat.refresh_grid()
    # copy of a known size (2^28 floats = 1 GiB read + 1 GiB written): the counter scale
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    _hip.call("copy", a.dtype, P(a), P(b), a.numel(), s)
    torch.cuda.synchronize()
    del a, b
    return out


def measured_traffic(npx, tile, passes, rule, multi, stream=False):
    """(HBM bytes per full-batch launch of the dominant kernel, where they come
    from): the committed counter runs (FETCH_SIZE and WRITE_SIZE, each scaled
    by what it reads for a copy of known size), when they were taken on this
    workload; else (None, None)."""
    if passes == 2:
        name = "tv_dual_kernel"
    elif multi > 1:
        name = ("tv_stream_kernel_k%d" if stream else "tv_multi_kernel_k%d") % multi
    else:
        name = "tv_iter_kernel"
    name += "_accel" if rule == "accel" else ""
    path = os.path.join(ROOT, "profiles", "%s_tvl2_256x1024_counters.json" % name)
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        m = json.load(fh)
    if m["workload"]["pixels"] != npx or list(m["workload"]["tile"]) != list(tile):
        return None, None
    kib = 1024.0
    fetch = m["fetch_kib"] * (m["copy_known_bytes_each_way"] / (m["copy_fetch_kib"] * kib))
    write = m["write_kib"] * (m["copy_known_bytes_each_way"] / (m["copy_write_kib"] * kib))
    return (fetch + write) * kib, ("not measured in this run: FETCH_SIZE + WRITE_SIZE of an earlier rocprofv3 "
                                   "--pmc run of this kernel and workload, read from %s"
                                   % os.path.relpath(path, ROOT))


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


def plain_same_rule(bat, specs, n, max_iters, args):
    """The batched solver's own step rule and kernels on one instance at a time
    (TVBatch of a single image, cold start): the same instances as the plain
    generic leg, so the ratio to it isolates what batching adds."""
    import torch
    from opymize_amd.solvers.tv import TVBatch
    f_all = bat.unpack(bat.f, "x")
    order = np.argsort([s["lam"] for s in specs])
    pick = [int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, n)]
    solved, iters, t_tot = 0, [], 0.0
    for b in pick:
        H, W = bat.shapes[b]
        one = TVBatch([f_all[b].reshape(H, W)], [specs[b]["lam"]], "l2", precision="float", tile=tuple(args.tile),
                      passes=args.passes, multi=min(bat.multi, 4), multi_kernel=bat.multi_kernel,
                      stream_rows=bat.stream_rows)
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = one.solve(steps=args.step_rule, step_bound=0.124, step_factor=args.step_factor, max_iters=max_iters,
                        block_iters=args.block_iters, tol=1e-5, certify=True, restart=not args.no_restart)
        torch.cuda.synchronize()
        t_tot += time.perf_counter() - t
        solved += res[0]["status"] == 1
        iters.append(int(res[0]["iter"]))
        del one
    return {"value": solved / t_tot, "instances": pick, "certified": int(solved), "iters": iters,
            "seconds": t_tot}


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _pool_worker(a):
    """One process of the CPU pool: the oracle on its own instance for a fixed
    number of iterations; returns (iterations, seconds)."""
    f, H, W, lam, rule, n_it = a
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import pdhg as opd
    prob = opd.tv_problem(f.reshape(-1, 1), (H, W), lam)
    t = time.perf_counter()
    opd.solve(prob, steps=rule, step_bound=0.124, term_maxiter=n_it, granularity=10 ** 9)
    return n_it, time.perf_counter() - t


def cpu_pool(bat, specs, procs, per_it_1core, budget_s, mean_iters, rule):
    """The oracle on `procs` processes at once, each on a different instance of
    the batch, for about budget_s seconds: aggregate iterations per second over
    the mean iterations per instance of the GPU run."""
    import multiprocessing as mp
    f_all = bat.unpack(bat.f, "x")
    n_it = max(2, int(budget_s / max(per_it_1core, 1e-4)))
    jobs = [(f_all[b], bat.shapes[b][0], bat.shapes[b][1], specs[b]["lam"], rule, n_it) for b in range(procs)]
    ctx = mp.get_context("spawn")
    t = time.perf_counter()
    with ctx.Pool(procs) as pool:
        out = pool.map(_pool_worker, jobs)
    wall = time.perf_counter() - t
    rate = sum(o[0] for o in out) / wall
    return {"processes": procs, "value": rate / mean_iters, "unit": "solves/s",
            "sample": "%d processes x %d oracle iterations (one instance each, 1 thread each) in %.1fs wall "
                      "(process start included), scaled to the mean %d iterations per instance of the GPU run"
                      % (procs, n_it, wall, mean_iters)}


def cpu_baseline(bat, specs, budget_s, gpu_iters, rule):
    """Oracle (float64 numpy/scipy, one thread, same step rule) on instance 0:
    seconds per iteration over a bounded sample, scaled by the mean
    iterations per instance of the GPU run (gpu_iters)."""
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
    model, ncpu = _cpu_info()
    return {"value": 1.0 / (per_it * gpu_iters), "unit": "solves/s", "cores": 1, "kind": "port",
            "cpu_model": model, "cpu_count": ncpu, "seconds_per_iteration": per_it,
            "sample": "oracle (float64 numpy/scipy, %s steps, 1 thread) on instance 0 (%dx%d): %d iterations "
                      "in %.1fs, scaled to the mean %d iterations per instance of the GPU run"
                      % (rule, H, W, n_it, dt, gpu_iters)}


LIFT_SIDE, LIFT_LABELS, LIFT_LAM, LIFT_SEED = 512, 16, 0.25, 4
LIFT_OMEGA = 6.0  # primal weight of the batched lifting solve (1 = the reference's preconditioner)


def lifting_kernels(bat, A, N, k, reps):
    """Eager full-batch launches of the fused lifting passes and, for scale,
    of the generic chain's per-pixel product on the matrix cores (r (N, k) ->
    z (N, 2k), the MatrixMultR of one instance), timed with HIP events on
    their stream.  Algorithmic bytes of a fused iteration: the primal pass
    reads U, r, p (2), z (2) and writes U, x_bar's U, r, x_bar's r: 10
    planes of k values per pixel; the dual pass reads x_bar (2), p (2), z (2)
    and f and writes p (2), z (2): 10 planes plus f."""
    import torch
    from opymize_amd import _hip
    stream = torch.cuda.current_stream()
    bat.reset()
    bat.refresh_grid()
    Pt, s = _hip.ptr, _hip.stream()
    ctl = lambda: (Pt(bat.insts), Pt(bat.active_inst), Pt(bat.act_off), Pt(bat.nact), bat.kmax, bat.grid, s)
    npl = sum(h * w * kk for (h, w), kk in zip(bat.shapes, bat.K))
    npx = sum(h * w for h, w in bat.shapes)
    es = 4 if bat.dtype == torch.float32 else 8
    mm = A.blocks[1][1]
    r = torch.rand(N * k, dtype=torch.float32, device="cuda")
    z = torch.zeros(N * 2 * k, dtype=torch.float32, device="cuda")
    cases = {
        "lift_primal": (lambda: bat._call("opy_lift_primal" + bat._sfx(), Pt(bat.U), Pt(bat.r), Pt(bat.Ub),
                                          Pt(bat.rb), Pt(bat.p), Pt(bat.z), *ctl()), 10 * npl * es),
        "lift_dual": (lambda: bat._call("opy_lift_dual" + bat._sfx(), Pt(bat.Ub), Pt(bat.rb), Pt(bat.p), Pt(bat.z),
                                        Pt(bat.f), *ctl()), 10 * npl * es + npx * es),
        "matmul_r_mfma (one instance, generic chain)": (lambda: mm(r, z), N * 3 * k * 4),
    }
    out = {}
    for name, (fn, nbytes) in cases.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    _hip.call("copy", a.dtype, _hip.ptr(a), _hip.ptr(b), a.numel(), _hip.stream())
    torch.cuda.synchronize()
    del a, b
    bat.reset()
    return out


def lifting_main(args, emit=True):
    """The multi-label lifting workload: a batch of 512x512, L = 16 instances
    solved together to relative gap 1e-5 (FP32 iterates, FP64 gap on the GPU)
    by the fused batched kernels with preconditioned steps and primal weight
    LIFT_OMEGA, against the plain single-instance drop-in PDHG (the generic
    kernel chain, the reference's preconditioner)."""
    import torch
    from opymize_amd import lifting
    from opymize_amd.solvers import PDHG
    from opymize_amd.solvers.lift import LiftBatch
    from opymize_amd.tools import synth
    n, L, lam = LIFT_SIDE, LIFT_LABELS, LIFT_LAM
    B = args.lift_batch
    imgs = [synth.rects(n, n, 1, seed=LIFT_SEED + b).reshape(n, n) for b in range(B)]
    N, k = n * n, L - 1
    bat = LiftBatch(imgs, L, lam, precision="float", omega=LIFT_OMEGA)
    G, F, A, gam = lifting.sublabel_tv(imgs[0], L, lam)
    kt = lifting_kernels(bat, A, N, k, args.reps)
    log("lifting kernels: %s" % ", ".join("%s %.4f ms %.0f GB/s" % (a, v["ms"], v["GBs"]) for a, v in kt.items()))
    if args.roofline_only:
        return

    def one_step():
        bat.reset()
        return bat.solve(max_iters=args.lifting_cap, block_iters=100, check_every=10, tol=1e-5)

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    certified, iters = 0, []
    for _ in range(args.steps):
        res = one_step()
        certified += sum(r["status"] == 1 for r in res)
        iters += [r["iter"] for r in res]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = certified / elapsed
    log("lifting batch: %d/%d certified in %.2fs, iterations %s" % (certified, args.steps * B, elapsed, iters))

    # the plain single-instance drop-in path (generic kernel chain) on instance 0: seconds per iteration
    # over a bounded sample, times the iterations of its rule on that instance (the fused kernels give the
    # same iterates up to rounding, so their count is measured there, in one single-instance solve)
    def generic_rate(weight, nit):
        s = PDHG(G, F, A)
        s.solve(steps="precond", precision="float", term_maxiter=200, granularity=200, term_pd_gap=1e-30,
                fused=False, precond_weight=weight)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s = PDHG(G, F, A)
        s.solve(steps="precond", precision="float", term_maxiter=nit, granularity=nit, term_pd_gap=1e-30,
                fused=False, precond_weight=weight)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / nit

    def single_iters(weight, cap):
        one = LiftBatch(imgs[:1], L, lam, precision="float", omega=weight)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = one.solve(max_iters=cap, block_iters=100, check_every=10, tol=1e-5)[0]
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    plain = None
    if args.plain > 0:
        r_ref, t_ref = single_iters(1.0, args.lifting_plain_cap)
        r_w, t_w = single_iters(LIFT_OMEGA, args.lifting_cap)
        spi = generic_rate(1.0, 3000)
        plain = {"solves_per_sec": (r_ref["status"] == 1) / (spi * r_ref["iter"]),
                 "seconds_per_iteration": spi, "iterations": r_ref["iter"], "certified": r_ref["status"] == 1,
                 "how": "generic chain, reference preconditioner: 3000 timed iterations on instance 0 x the "
                        "iterations that rule needs there (fused single-instance solve, same iterates up to "
                        "rounding)",
                 "single_instance_fused_reference_rule_solves_per_sec": (r_ref["status"] == 1) / t_ref,
                 "single_instance_fused_same_rule_solves_per_sec": (r_w["status"] == 1) / t_w,
                 "single_instance_fused_same_rule_iterations": r_w["iter"]}
        log("lifting plain: %s" % plain)
    cpu = None
    if not args.no_cpu:
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        from oracle import lifting as olf, pdhg as opd
        prob, _ = olf.problem(imgs[0], L, lam)
        n_it = 3
        t = time.perf_counter()
        opd.solve(prob, steps="precond", term_maxiter=n_it, granularity=10 ** 9, term_pd_gap=1e-30,
                  precond_weight=LIFT_OMEGA)
        dt = time.perf_counter() - t
        model, ncpu = _cpu_info()
        it0 = int(round(np.mean(iters)))
        cpu = {"value": 1.0 / (dt / n_it * it0), "unit": "solves/s", "cores": 1, "kind": "port",
               "cpu_model": model, "cpu_count": ncpu,
               "sample": "oracle (float64 numpy/scipy, 1 thread): %d preconditioned iterations of one 512x512, "
                         "L=16 instance in %.1fs (step set-up included), scaled to the mean %d iterations per "
                         "instance of the GPU batch" % (n_it, dt, it0)}
    prim, dual = kt["lift_primal"], kt["lift_dual"]
    it_ms = prim["ms"] + dual["ms"]
    it_bytes = prim["bytes"] + dual["bytes"]
    achieved = it_bytes / (it_ms * 1e-3) / 1e9
    line = {
        "metric": "solves/sec to rel PD-gap 1e-5: batched+warm-start vs plain PDHG; %HBM peak",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "multi-label lifting, %d instances of %dx%d (rectangle generator, seeds %d..), "
                               "L=%d labels, (s - f)^2 data pieces as quadratic-epigraph constraints, TV weight "
                               "%g, preconditioned steps with primal weight %g, rel gap 1e-5 FP64 on the GPU, cap "
                               "%d iterations, cold start"
                               % (B, n, n, LIFT_SEED, L, lam, LIFT_OMEGA, args.lifting_cap),
                   "path": "fused batched kernels (LiftBatch)", "certified": int(certified),
                   "instances": int(args.steps * B), "iterations": iters,
                   "plain_single_instance": plain,
                   "ratio_vs_plain": (value / plain["solves_per_sec"]) if plain and plain["solves_per_sec"] else None,
                   "warm_start": "none: a start lifted from the ROF solution (primal, or primal and dual) needed "
                                 "more iterations than the cold start",
                   "kernels": kt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "traffic_source": None,
                     "kernel": "lift_primal_kernel<float,16> + lift_dual_kernel<float,16>: one iteration over the "
                               "batch, %d B algorithmic (20 planes of k=%d FP32 values per pixel, and f)"
                               % (it_bytes, k),
                     "avg_ms": it_ms},
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def lifting_summary(args):
    """The lifting workload's figures for the default line: one timed batch
    solve (no warm-up solve; its graphs are captured during it), the plain
    legs and the fused kernels' roofline, no CPU leg."""
    import copy
    a = copy.copy(args)
    a.warmup, a.steps, a.no_cpu, a.roofline_only, a.reps = 0, 1, True, False, min(args.reps, 10)
    line = lifting_main(a, emit=False)
    c = line["config"]
    return {"workload": c["workload"], "solves_per_sec": line["value"], "certified": c["certified"],
            "instances": c["instances"], "seconds": line["ms_per_step"] / 1e3, "iterations": c["iterations"],
            "plain_single_instance": c["plain_single_instance"], "ratio_vs_plain": c["ratio_vs_plain"],
            "roofline": {k: line["roofline"][k] for k in ("achieved", "peak", "unit", "frac", "kernel", "avg_ms")},
            "timed_first_solve": True}


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
    ap.add_argument("--multi", type=int, default=8, choices=[1, 2, 3, 4, 5, 6, 7, 8],
                    help="iterations per launch (accelerated steps; adaptive steps take 1)")
    ap.add_argument("--passes", type=int, default=1, choices=[1, 2])
    ap.add_argument("--multi-kernel", default="auto", choices=["auto", "stream", "tile"],
                    help="work split of the K-iteration launches (auto: streaming strips, tiles for precond steps)")
    ap.add_argument("--stream-rows", type=int, default=128, help="rows per unit of the streaming kernel")
    ap.add_argument("--warm", action="store_true",
                    help="time the learned warm start (the network's prediction of x, y) instead of the cold start")
    ap.add_argument("--cold", action="store_true", help="start from x = y = 0 (the default)")
    ap.add_argument("--workload", default="tv", choices=["tv", "lifting"],
                    help="tv: the batched TV-L2 solve; lifting: the 512x512, L=16 multi-label lifting model")
    ap.add_argument("--lifting-cap", type=int, default=200000, help="iteration cap of the lifting solve")
    ap.add_argument("--lift-batch", type=int, default=8, help="instances of the lifting batch")
    ap.add_argument("--no-lifting", action="store_true",
                    help="leave the lifting workload's summary out of the default line")
    ap.add_argument("--lifting-plain-cap", type=int, default=400000,
                    help="iteration cap of the reference-rule single-instance lifting solve")
    ap.add_argument("--plain-same-rule", type=int, default=6,
                    help="instances for the single-instance fused baseline with the batched step rule")
    ap.add_argument("--pool", type=int, default=16, help="processes of the CPU pool baseline (0: skip)")
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
    if args.workload == "lifting":
        return lifting_main(args)
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

    warm = bool(args.warm) and not args.cold

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
                 "iterations_median": float(np.median([r["iter"] for r in res])),
                 "iterations_max": int(max(r["iter"] for r in res)),
                 "fallbacks": len(bat.fallbacks) if not warm else None}
        other["ratio_to_timed"] = other["solves_per_sec"] / value if value > 0 else None
        log("batched (%s start, for comparison): %d/%d certified in %.2fs" % (other["start"], c, args.batch, dt))

    kt = roofline_phase(bat, args.reps, args.step_rule)
    stream_used = bat.multi > 1 and bat._use_stream(bat._cfg(args.step_rule))
    stream_rows = bat.stream_rows
    variant = "adaptive" if args.step_rule == "adaptive" else "accelerated steps"
    k_launch = 1
    if bat.passes == 2:
        kname, kms, kbytes = "dual", kt["dual"], DUAL_BYTES_PER_PX
        kdesc = "tv_dual_kernel<float,1,%d>, %s (second of two passes)" % (args.tile[1], variant)
    elif "multi" in kt:
        k_launch = bat.multi
        kname, kms, kbytes = "multi", kt["multi"], MULTI_BYTES_PER_PX
        if stream_used:
            kdesc = ("tv_stream_kernel<float,1,%d>, %s (%d iterations per launch, one wavefront per %d x %d unit, "
                     "the iterations pipelined down its rows)" % (k_launch, variant, k_launch, 64 - 2 * k_launch,
                                                                  stream_rows))
        else:
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
    same = plain_same_rule(bat, specs, args.plain_same_rule, args.max_iters, args) \
        if args.plain_same_rule > 0 else None
    if same:
        log("single-instance fused, same rule: %d/%d certified, %.2f solves/s, iterations %s"
            % (same["certified"], len(same["instances"]), same["value"], same["iters"]))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bat, specs, args.cpu_budget, int(round(iters.mean())), args.step_rule)
        log("cpu oracle: %s" % cpu["sample"])
        if args.pool > 0:
            cpu["pool"] = cpu_pool(bat, specs, min(args.pool, args.batch), cpu["seconds_per_iteration"],
                                   args.cpu_budget, int(round(iters.mean())), args.step_rule)
            log("cpu pool: %s" % cpu["pool"]["sample"])

    lift = None
    if not args.no_lifting:
        del bat
        torch.cuda.empty_cache()
        lift = lifting_summary(args)
        log("lifting summary: %d/%d certified, %.3f solves/s, %.1fx the plain chain, fused iteration at %.0f GB/s"
            % (lift["certified"], lift["instances"], lift["solves_per_sec"], lift["ratio_vs_plain"] or 0.0,
               lift["roofline"]["achieved"]))

    from opymize_amd.solvers import tv as tvmod
    traffic = measured_traffic(npx, args.tile, args.passes, args.step_rule, k_launch, stream=stream_used)
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
                   "multi_kernel": ("stream" if stream_used else "tile") if k_launch > 1 else None,
                   "stream_rows": stream_rows if stream_used else None,
                   "certified": int(certified), "instances": int(args.steps * args.batch),
                   "iterations_median": float(np.median(iters)), "iterations_max": int(iters.max()),
                   "plain_single_instance_solves_per_sec": plain["value"] if plain else None,
                   "plain_instances": len(plain["instances"]) if plain else 0,
                   "plain_certified": plain["certified"] if plain else 0,
                   "ratio_vs_plain": (value / plain["value"]) if plain and plain["value"] > 0 else None,
                   "single_instance_same_rule_solves_per_sec": same["value"] if same else None,
                   "single_instance_same_rule_certified": same["certified"] if same else 0,
                   "ratio_vs_single_instance_same_rule":
                       (value / same["value"]) if same and same["value"] > 0 else None,
                   "lifting": lift},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic[0], "traffic_source": traffic[1],
                     "kernel": "%s: %d B/pixel algorithmic x %d pixels per launch" % (kdesc, kbytes, npx),
                     "avg_ms": kms, "ms_per_iteration": kms / k_launch,
                     "launch_ms": {k: v for k, v in kt.items() if not k.startswith("grid")},
                     "blocks_per_launch": kt["grid"]},
        "cpu_baseline": cpu,
    }
    if "multi_k4" in kt:
        # the same kernel at 4 iterations per launch (same bytes per launch):
        # beyond 4 the launch is bound by instruction issue, not HBM, and it is
        # the time per iteration the solve needs
        a4 = kbytes * npx / (kt["multi_k4"] * 1e-3) / 1e9
        line["roofline"]["note"] = ("%d iterations per launch: bound by instruction issue (VALU; SQ_INSTS_VALU in "
                                    "the committed profile of this kernel), %.3f ms per iteration, %.1f TB/s of "
                                    "single-iteration traffic (28 B/pixel per iteration); the same kernel at 4 per "
                                    "launch is HBM-bound, see roofline_k4"
                                    % (k_launch, kms / k_launch, achieved * k_launch / 1e3))
        line["roofline_k4"] = {"bound": "hbm", "achieved": a4, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": a4 / HBM_PEAK_GBS, "avg_ms": kt["multi_k4"],
                               "ms_per_iteration": kt["multi_k4"] / 4, "blocks_per_launch": kt["grid_k4"],
                               "kernel": "the same streaming kernel with 4 iterations per launch (%d x %d units), "
                                         "%d B/pixel algorithmic x %d pixels per launch"
                                         % (56, stream_rows, kbytes, npx)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
