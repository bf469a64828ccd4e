"""The benchmark workloads at their full sizes.

- ROF 256x256 FP64 (lambda 8, adaptive steps, 1000 iterations): the oracle
  runs at full size; iterates and the final gap agree to 1e-9 (max abs /
  relative to the primal value) on the generic and the fused path.
- Vector TV 2048x2048x3 FP32 (Frobenius ball, channel-correlated noise,
  lambda 8, adaptive steps, 5000 iterations): the first 3 iterations against
  the FP64 oracle at full size (FP32 5e-4 max abs); the 5000-iteration FP32
  run against the same run in FP64 on the GPU through the strong-convexity
  certificate: 0.5 lambda |x - x*|^2 <= gap for both points, so
  |x32 - x64| <= sqrt(2 gap32 / lambda) + sqrt(2 gap64 / lambda).
- Batched uneven TV-L1 + TV-L2 (512 instances, sides 64..1024, seed 3):
  100 iterations of the whole batch in blocks of 50; the 6 smallest
  instances against their oracle runs (FP32 5e-4), every gap finite.
- The learned-warm-start batch (256 x 1024^2 TV-L2): 48 iterations with
  4 iterations per launch against single-iteration launches (1e-6), and
  instance 0 against the oracle (5e-4); and its smallest-lambda instance in
  FP64 with restarts of the accelerated schedule, 240 iterations against the
  oracle's restart rule (same restarts, 1e-9).
"""
import numpy as np
import pytest

from oracle import pdhg as opd

pytestmark = pytest.mark.gpu


def _tv_model(f, dims, lam):
    from opymize_amd.functionals import SSD, L1Norms
    from opymize_amd.linear import GradientOp
    N, C = f.shape
    return SSD(f, vol=lam * np.ones(C)), L1Norms(N, (C, len(dims)), lbd=1.0), GradientOp(dims, C, scheme="forward")


@pytest.mark.parametrize("fused", [False, True])
def test_rof_256_fp64_adaptive_1000(gpu, fused):
    from opymize_amd.solvers import PDHG
    from opymize_amd.tools import synth
    H = W = 256
    f = synth.rects(H, W, 1, seed=1).reshape(-1, 1)
    G, F, A = _tv_model(f, (H, W), 8.0)
    k = 1000
    s = PDHG(G, F, A)
    info = s.solve(steps="adaptive", precision="double", step_bound=0.124, term_maxiter=k, granularity=k,
                   term_pd_gap=1e-30, fused=fused)
    x, y = s.state
    xo, yo, io, _ = opd.solve(opd.tv_problem(f, (H, W), 8.0), steps="adaptive", step_bound=0.124, term_maxiter=k,
                              granularity=k, term_pd_gap=1e-30)
    print("ROF 256^2 FP64, 1000 iterations, fused=%s: max |x - x_oracle| = %.2e, max |y - y_oracle| = %.2e, "
          "gap %.6e (oracle %.6e)" % (fused, np.abs(x - xo).max(), np.abs(y - yo).max(), info["gap"], io["gap"]))
    assert np.abs(x - xo).max() < 1e-9 and np.abs(y - yo).max() < 1e-9
    assert abs(info["gap"] - io["gap"]) <= 1e-9 * abs(io["objp"])
    assert abs(s.itervars["tauk"] - io["tau"]) <= 1e-9 * io["tau"]


def test_vector_tv_2048x3_fp32(gpu):
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    H = W = 2048
    lam = 8.0
    img = synth.rects(H, W, 3, seed=2, corr=0.5)
    # first iterations against the oracle at full size
    k0 = 3
    bat = TVBatch([img], lam, "l2", precision="float", tile=(64, 4))
    bat.solve(steps="adaptive", step_bound=0.124, max_iters=k0, block_iters=k0, tol=1e-30)
    xs, ys = bat.solution()
    xo, yo, io, _ = opd.solve(opd.tv_problem(img.reshape(-1, 3), (H, W), lam), steps="adaptive", step_bound=0.124,
                              term_maxiter=k0, granularity=k0, term_pd_gap=1e-30)
    ex, ey = np.abs(xs[0] - xo.ravel()).max(), np.abs(ys[0] - yo.ravel()).max()
    print("vector TV 2048^2x3 FP32, %d iterations: max |x - x_oracle| = %.2e, max |y - y_oracle| = %.2e" % (k0, ex, ey))
    assert ex < 5e-4 and ey < 5e-4
    del bat
    # the full run, FP32 against FP64
    k = 5000
    out = {}
    for prec in ("float", "double"):
        b = TVBatch([img], lam, "l2", precision=prec, tile=(64, 4))
        r = b.solve(steps="adaptive", step_bound=0.124, max_iters=k, block_iters=50, tol=1e-30)[0]
        out[prec] = (b.solution()[0][0], r)
        del b
    (x32, r32), (x64, r64) = out["float"], out["double"]
    d = np.linalg.norm(x32 - x64)
    bound = np.sqrt(2 * max(r32["gap"], 0) / lam) + np.sqrt(2 * max(r64["gap"], 0) / lam)
    print("vector TV 2048^2x3, %d iterations: relgap FP32 %.3e, FP64 %.3e; |x32 - x64| = %.3e (bound %.3e), "
          "max abs %.2e" % (k, r32["relgap"], r64["relgap"], d, bound, np.abs(x32 - x64).max()))
    assert r32["iter"] == k and r64["iter"] == k
    assert np.isfinite(r32["gap"]) and abs(r32["relgap"]) < 1e-2 and abs(r64["relgap"]) < 1e-2
    assert d <= bound * (1 + 1e-6) + 1e-6


def test_batched_uneven_512(gpu):
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    specs = synth.batch_uneven(512, seed=3)
    imgs = synth.images(specs)
    npx = sum(im.size for im in imgs)
    assert 150e6 < npx < 250e6
    lams = [s["lam"] for s in specs]
    data = [s["data"] for s in specs]
    assert 0.4 < np.mean([d == "l1" for d in data]) < 0.6
    bat = TVBatch(imgs, lams, data, precision="float", tile=(64, 8))
    k = 100
    res = bat.solve(steps="adaptive", step_bound=0.124, max_iters=k, block_iters=50, tol=1e-30)
    assert all(np.isfinite(r["gap"]) and r["iter"] == k for r in res)
    order = np.argsort([im.size for im in imgs])[:6]
    xs = bat.unpack(bat.x[0], "x")
    worst = 0.0
    for b in order:
        H, W = imgs[b].shape
        xo, _, io, _ = opd.solve(opd.tv_problem(imgs[b].reshape(-1, 1), (H, W), lams[b], data[b]), steps="adaptive",
                                 step_bound=0.124, term_maxiter=k, granularity=k, term_pd_gap=1e-30)
        e = np.abs(xs[b] - xo).max()
        worst = max(worst, e)
        assert e < 5e-4, (b, H, W, data[b])
    print("batched uneven: %d instances, %.1f M pixels; 6 smallest after %d iterations: max |x - x_oracle| = %.2e"
          % (len(imgs), npx / 1e6, k, worst))


def test_batched_uneven_512_certifies_every_instance(gpu):
    """The uneven TV-L1 + TV-L2 batch (512 instances, sides 64..1024, seed 3) on
    the path the design recommends: accelerated steps with restarts for the
    TV-L2 instances, the TV-L1 instances on constant steps from their own
    split (step_factor_l1 = 1/256), tiles and iterations per launch from the
    autotuner, FP32, checks every 48 iterations.  Every instance is certified
    at relative gap 1e-5 within 20000 iterations (the TV-L1 instances with
    lambda close to 1 need 10000-15000), and the six smallest instances agree
    with the oracle run for the same number of iterations with the same rule."""
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.solvers import autotune
    from opymize_amd.tools import synth
    specs = synth.batch_uneven(512, seed=3)
    imgs = synth.images(specs)
    lams = [s["lam"] for s in specs]
    data = [s["data"] for s in specs]
    bat = TVBatch(imgs, lams, data, precision="float", tile=(64, 6), multi=4)
    best = autotune.tune(bat, steps="accel")
    print("uneven batch: tuned %s" % (best,), flush=True)
    import time
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, step_factor_l1=1.0 / 256, max_iters=20000,
                    block_iters=48, tol=1e-5, certify=True)
    torch.cuda.synchronize()
    print("uneven batch: solved in %.2fs" % (time.perf_counter() - t0))
    assert all(np.isfinite(r["gap"]) for r in res)
    nl1 = sum(d == "l1" for d in data)
    cert = sum(r["status"] == 1 for r in res)
    print("uneven batch: %d / 512 certified (TV-L1 %d, TV-L2 %d), iterations max %d"
          % (cert, sum(r["status"] == 1 for r, d in zip(res, data) if d == "l1"),
             sum(r["status"] == 1 for r, d in zip(res, data) if d == "l2"), max(r["iter"] for r in res)))
    assert cert == 512 and nl1 > 200
    assert all(abs(r["relgap_cert"]) < 1e-5 for r in res)
    xs, _ = bat.solution()
    for b in np.argsort([im.size for im in imgs])[:6]:
        H, W = imgs[b].shape
        prob = opd.tv_problem(imgs[b].reshape(-1, 1), (H, W), lams[b], data[b])
        it = res[b]["iter"]
        xo, _, io, _ = opd.solve(prob, steps="accel", step_bound=0.124,
                                 step_factor=(1.0 / 256) if data[b] == "l1" else 4.0, term_maxiter=it,
                                 granularity=48, term_pd_gap=1e-30, restart=(0.5, 96))
        e = np.abs(xs[b] - xo).max()
        print("instance %d (%dx%d, %s, lambda %.2f): %d iterations, max |x - x_oracle| = %.2e"
              % (b, H, W, data[b], lams[b], it, e))
        assert io["iter"] == it
        assert e < 2e-3, (b, e)


def test_warm_start_batch_256x1024(gpu):
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    specs = synth.batch_square(256, side=1024, seed=5)
    imgs = synth.images(specs)
    lams = [s["lam"] for s in specs]
    k = 48
    out = {}
    for m in (1, 4):
        bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, 8), multi=m)
        bat.solve(steps="accel", step_bound=0.124, max_iters=k, block_iters=k, tol=1e-30)
        out[m] = (bat.x[0].cpu().numpy(), bat.y[0].cpu().numpy(), bat.unpack(bat.x[0], "x")[0])
        del bat
    assert np.abs(out[1][0] - out[4][0]).max() <= 1e-6 and np.abs(out[1][1] - out[4][1]).max() <= 1e-6
    xo, _, io, _ = opd.solve(opd.tv_problem(imgs[0].reshape(-1, 1), (1024, 1024), lams[0]), steps="accel",
                             step_bound=0.124, term_maxiter=k, granularity=k, term_pd_gap=1e-30)
    e = np.abs(out[4][2] - xo).max()
    print("256 x 1024^2, %d accelerated iterations, 4 per launch: max |x - x_oracle| on instance 0 = %.2e" % (k, e))
    assert e < 5e-4


def test_restarts_at_full_size_1024(gpu):
    """The bench's solver on one 1024^2 instance of its batch (the smallest
    lambda) in FP64: 4 iterations per launch on 64 x 24 tiles, accelerated
    steps with restarts (tests every 24 iterations, decay 0.5, at least 48
    iterations apart), 240 iterations; the same restarts and iterates as the
    oracle's rule at full size (1e-9)."""
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    specs = synth.batch_square(256, side=1024, seed=5)
    b = int(np.argmin([s["lam"] for s in specs]))
    img = synth.images([specs[b]])[0]
    lam = specs[b]["lam"]
    k = 240
    bat = TVBatch([img], [lam], "l2", precision="double", tile=(64, 6), multi=4)
    res = bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, max_iters=k, block_iters=24, tol=1e-30,
                    restart=(0.5, 48))
    xs, ys = bat.solution()
    xo, yo, io, _ = opd.solve(opd.tv_problem(img.reshape(-1, 1), (1024, 1024), lam), steps="accel",
                              step_bound=0.124, step_factor=4.0, term_maxiter=k, granularity=24, term_pd_gap=1e-30,
                              restart=(0.5, 48))
    e = max(np.abs(xs[0] - xo).max(), np.abs(ys[0] - yo).max())
    print("1024^2, lambda %.3f, %d iterations: restarts GPU %d, oracle %s; max |GPU - oracle| = %.2e"
          % (lam, k, res[0]["restarts"], io["restarts"], e))
    assert res[0]["restarts"] == len(io["restarts"]) and res[0]["restarts"] > 0
    assert e < 1e-9
