"""Batched fused TV PDHG: every instance of an uneven batch against its own
oracle run.  Tolerances as in test_gpu_pdhg.py (FP64 1e-9, FP32 5e-4 max abs)."""
import numpy as np
import pytest

from oracle import pdhg as opd

pytestmark = pytest.mark.gpu


def rects(H, W, seed, nrect=6):
    rng = np.random.default_rng(seed)
    img = np.full((H, W), 0.5)
    for _ in range(nrect):
        y0, y1 = sorted(rng.integers(0, H, 2))
        x0, x1 = sorted(rng.integers(0, W, 2))
        img[y0:y1 + 1, x0:x1 + 1] = rng.random()
    return img + 0.1 * rng.standard_normal((H, W))


SHAPES = [(17, 30), (40, 23), (5, 64), (64, 5), (1, 33), (70, 70)]


def make_batch(precision, tile=(64, 2), passes=1):
    from opymize_amd.solvers.tv import TVBatch
    imgs = [rects(H, W, i) for i, (H, W) in enumerate(SHAPES)]
    lams = [2.0, 7.5, 1.0, 15.0, 4.0, 3.3]
    data = ["l2", "l1", "l2", "l1", "l2", "l2"]
    return TVBatch(imgs, lams, data, precision=precision, tile=tile, passes=passes), imgs, lams, data


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("steps", ["adaptive", "const", "precond"])
@pytest.mark.parametrize("passes,tile", [(2, (64, 2)), (2, (32, 1)), (2, (16, 4)), (1, (64, 1)), (1, (64, 4)),
                                          (1, (64, 8))])
def test_batch_iterates_match_oracle(gpu, precision, steps, passes, tile):
    bat, imgs, lams, data = make_batch(precision, tile, passes)
    k = 40
    bat.solve(steps=steps, step_bound=0.124, max_iters=k, block_iters=20, tol=1e-30)
    xs, ys = bat.solution()
    res = bat.results()
    tol = 1e-9 if precision == "double" else 5e-4
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps=steps, step_bound=0.124, term_maxiter=k, granularity=k,
                                  term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b
        assert np.abs(ys[b] - yo).max() < tol, b
        rel = 1e-9 if precision == "double" else 1e-4
        assert abs(res[b]["gap"] - io["gap"]) <= rel * max(1.0, abs(io["objp"])), b
        assert res[b]["iter"] == k and res[b]["status"] == 0
    if steps == "adaptive":
        st = bat.steps_state()
        for b, im in enumerate(imgs):
            H, W = im.shape
            prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
            _, _, io, _ = opd.solve(prob, steps=steps, step_bound=0.124, term_maxiter=k, granularity=k,
                                    term_pd_gap=1e-30)
            assert abs(st[b]["tau"] - io["tau"]) <= 1e-9 * io["tau"], b


def test_batch_converges_certified_and_drops_instances(gpu):
    bat, imgs, lams, data = make_batch("float")
    seen = []
    res = bat.solve(steps="adaptive", step_bound=0.124, max_iters=20000, block_iters=50, tol=1e-5,
                    callback=lambda e, it: seen.append(int(e.nact.item())))
    assert all(r["status"] == 1 for r in res), res
    assert all(abs(r["relgap_cert"]) < 1e-5 for r in res)
    # instances leave the running set as they converge
    assert seen == sorted(seen, reverse=True) and seen[-1] == 0
    xs, _ = bat.solution()
    for b, im in enumerate(imgs):
        if data[b] != "l2":
            continue
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b])
        xo, _, io, _ = opd.solve(prob, steps="adaptive", step_bound=0.124, term_maxiter=30000, granularity=50,
                                 term_pd_gap=1e-10)
        bound = np.sqrt(2 * max(res[b]["gap"], 0) / lams[b]) + np.sqrt(2 * max(io["gap"], 0) / lams[b])
        assert np.linalg.norm(xs[b] - xo) <= 1.01 * bound + 1e-6, b


def test_batch_graph_equals_eager(gpu):
    out = {}
    for g in (False, True):
        bat, *_ = make_batch("double")
        bat.solve(steps="adaptive", step_bound=0.124, max_iters=200, block_iters=20, tol=1e-30, use_graph=g)
        out[g] = bat.solution()[0]
    for a, b in zip(out[False], out[True]):
        np.testing.assert_array_equal(a, b)


def test_batch_warm_start_from_solution(gpu):
    """Continuing from a converged solution certifies at the first check."""
    bat, imgs, lams, data = make_batch("double")
    bat.solve(steps="adaptive", step_bound=0.124, max_iters=20000, tol=1e-7)
    x, y = bat.x[0].clone(), bat.y[0].clone()
    bat2, *_ = make_batch("double")
    bat2.set_start(x, y)
    res = bat2.solve(steps="adaptive", step_bound=0.124, max_iters=1000, tol=1e-6, check_start=True)
    assert all(r["status"] == 1 for r in res) and all(r["iter"] == 0 for r in res)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("steps", ["adaptive", "precond"])
@pytest.mark.parametrize("passes,tile", [(2, (128, 8)), (2, (256, 16)), (2, (64, 4)), (1, (64, 2)), (1, (64, 8))])
@pytest.mark.parametrize("C", [1, 2, 3, 4])
def test_batch_channels_and_wide_tiles(gpu, precision, steps, passes, tile, C):
    """Multi-channel batches (per-pixel vectors of C and 2C values) and tiles
    whose rows span several wavefronts, against the oracle after 30 iterations."""
    from opymize_amd.solvers.tv import TVBatch
    shapes = [(33, 150), (9, 300), (70, 41)]
    rng = np.random.default_rng(C)
    imgs = [np.stack([rects(H, W, 10 * C + i + k) for k in range(C)], axis=2) for i, (H, W) in enumerate(shapes)]
    lams = [3.0, 9.0, 1.5]
    data = ["l2", "l1", "l2"]
    if C == 4 and tile[1] == 16:
        # the shape is refused on the host (the entry points' own check: tests/test_abi.py)
        with pytest.raises(ValueError, match="register budget"):
            TVBatch(imgs, lams, data, precision=precision, tile=tile, passes=passes)
        return
    bat = TVBatch(imgs, lams, data, precision=precision, tile=tile, passes=passes)
    k = 30
    bat.solve(steps=steps, step_bound=0.124, max_iters=k, block_iters=10, tol=1e-30)
    xs, ys = bat.solution()
    tol = 1e-9 if precision == "double" else 5e-4
    for b, im in enumerate(imgs):
        H, W = im.shape[:2]
        prob = opd.tv_problem(im.reshape(-1, C), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps=steps, step_bound=0.124, term_maxiter=k, granularity=k,
                                  term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b
        assert np.abs(ys[b] - yo).max() < tol, b


@pytest.mark.parametrize("passes", [1, 2])
def test_autotune_picks_a_candidate_and_keeps_results(gpu, passes):
    """The tuner times the candidate tiles on the batch itself, restores the
    start point and leaves the winner set; iterates afterwards still match
    the oracle."""
    from opymize_amd.solvers import autotune
    bat, imgs, lams, data = make_batch("double", (64, 2), passes)
    autotune._cache.clear()
    best = autotune.tune(bat, iters=4, blocks=(10, 20))
    cands = autotune.ITER_ROWS_PER_THREAD if passes == 1 else autotune.ROWS_PER_THREAD
    assert best["tile"][1] in cands and (bat.tw, bat.rpt) == tuple(best["tile"])
    assert best["block_iters"] in (10, 20)
    k = 20
    bat.solve(steps="adaptive", step_bound=0.124, max_iters=k, block_iters=10, tol=1e-30)
    xs, _ = bat.solution()
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, _, _, _ = opd.solve(prob, steps="adaptive", step_bound=0.124, term_maxiter=k, granularity=k,
                                term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < 1e-9, b


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("passes,tile", [(1, (64, 8)), (1, (64, 4)), (2, (64, 2))])
def test_accelerated_steps_match_oracle(gpu, precision, passes, tile):
    """steps="accel": per-instance theta, tau, sigma of Chambolle & Pock's
    Algorithm 2 on the TV-L2 instances (the TV-L1 ones keep constant steps)."""
    bat, imgs, lams, data = make_batch(precision, tile, passes)
    k = 40
    bat.solve(steps="accel", step_bound=0.124, max_iters=k, block_iters=20, tol=1e-30)
    xs, ys = bat.solution()
    st = bat.steps_state()
    tol = 1e-9 if precision == "double" else 5e-4
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps="accel", step_bound=0.124, term_maxiter=k, granularity=k,
                                  term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b
        assert np.abs(ys[b] - yo).max() < tol, b
        assert abs(st[b]["tau"] - io["tau"]) <= 1e-12 * io["tau"], b
        assert abs(st[b]["sigma"] - io["sigma"]) <= 1e-12 * io["sigma"], b


def test_warm_start_fallback_restarts_cold(gpu):
    """fallback=True: an instance whose start is not finite is restarted
    cold (x = y = 0, fresh steps and restart record) at the first test and
    then certifies like the others; the instances with a usable start keep
    it.  FP64, accelerated steps with restarts and the TV-L1 split."""
    bat, imgs, lams, data = make_batch("double")
    ref, *_ = make_batch("double")
    ref.solve(steps="accel", step_bound=0.124, step_factor=4.0, step_factor_l1=1.0 / 256, max_iters=20000,
              block_iters=40, tol=1e-7)
    x0 = bat.pack([im for im in imgs], "x")  # x = f: a usable start
    bad = 2
    x0[bat.offs[bad]:bat.offs[bad] + imgs[bad].size] = np.nan
    bat.set_start(x0, None)
    res = bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, step_factor_l1=1.0 / 256, max_iters=20000,
                    block_iters=40, tol=1e-7, fallback=True)
    assert list(bat.fallbacks) == [bad]
    assert all(r["status"] == 1 for r in res), res
    xs, _ = bat.solution()
    xr, _ = ref.solution()
    rr = ref.results()
    for b, lam in enumerate(lams):
        if data[b] != "l2":
            continue
        bound = np.sqrt(2 * max(res[b]["gap"], 0) / lam) + np.sqrt(2 * max(rr[b]["gap"], 0) / lam)
        assert np.linalg.norm(xs[b] - xr[b]) <= 1.01 * bound + 1e-6, b
