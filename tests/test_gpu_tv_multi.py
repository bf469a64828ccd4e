"""Several fused TV iterations per launch (opy_tv_multi, TVBatch(multi=K,
multi_kernel="tile")): the iterates of K iterations on LDS-staged tiles
against the oracle and against K single-iteration launches (the streaming
kernel has its own tests in test_gpu_tv_stream.py).  Tolerances as in test_gpu_tv_batch.py
(FP64 1e-9, FP32 5e-4 max abs against the oracle); against the single-launch
path 1e-6 (FP32) and 1e-12 (FP64): the per-pixel arithmetic is the same, so
only the compiler's instruction selection can differ."""
import numpy as np
import pytest

from oracle import pdhg as opd

from test_gpu_tv_batch import SHAPES, make_batch, rects

pytestmark = pytest.mark.gpu


def lds_bytes(precision, C, rpt, k):
    """LDS of one block: y and xbar (3C values) over the tile and k pixels
    around it, rows rounded up to a multiple of 4, and a guard pixel on every
    side; and the k iterations' scalar steps (3 + 4C values each)."""
    rows = 4 * ((4 * rpt + 2 * k + 3) // 4) + 2
    size = 8 if precision == "double" else 4
    return rows * (64 + 2 * k + 2) * 3 * C * size + k * (3 + 4 * C) * size


def _batch(precision, tile, multi, shapes=SHAPES, C=1, seed=0, data=None, kernel="tile"):
    from opymize_amd.solvers.tv import TVBatch
    rng = np.random.default_rng(seed)
    if C == 1:
        imgs = [rects(H, W, seed + i) for i, (H, W) in enumerate(shapes)]
    else:
        imgs = [np.stack([rects(H, W, 10 * C + i + k) for k in range(C)], axis=2) for i, (H, W) in enumerate(shapes)]
    lams = list(np.round(np.exp(rng.uniform(0.0, np.log(20.0), len(shapes))), 3))
    if data is None:
        data = ["l2" if i % 3 != 1 else "l1" for i in range(len(shapes))]
    return (TVBatch(imgs, lams, data, precision=precision, tile=tile, multi=multi, multi_kernel=kernel), imgs, lams,
            data)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("steps", ["const", "precond", "accel"])
@pytest.mark.parametrize("multi,rpt", [(2, 8), (3, 4), (4, 2), (4, 8)])
def test_multi_matches_oracle(gpu, precision, steps, multi, rpt):
    """40 iterations in blocks of 20 (K=3: six launches of 3 and two single
    iterations per block) on the uneven batch, including 1-row, 5-row and
    5-column images."""
    from opymize_amd import _hip
    bat, imgs, lams, data = _batch(precision, (64, rpt), multi)
    k = 40
    if lds_bytes(precision, 1, rpt, multi) > 64 * 1024:  # FP64 with 8 rows per thread and K = 3 or 4
        with pytest.raises(_hip.HipError, match="LDS"):
            bat.solve(steps=steps, step_bound=0.124, max_iters=k, block_iters=20, tol=1e-30)
        return
    bat.solve(steps=steps, step_bound=0.124, max_iters=k, block_iters=20, tol=1e-30)
    xs, ys = bat.solution()
    res = bat.results()
    st = bat.steps_state()
    tol = 1e-9 if precision == "double" else 5e-4
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps=steps, step_bound=0.124, term_maxiter=k, granularity=k,
                                  term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b
        assert np.abs(ys[b] - yo).max() < tol, b
        assert res[b]["iter"] == k and res[b]["status"] == 0
        if steps == "accel":
            assert abs(st[b]["tau"] - io["tau"]) <= 1e-12 * io["tau"], b


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("multi,rpt", [(2, 8), (3, 4), (4, 4)])
@pytest.mark.parametrize("steps", ["accel", "precond"])
def test_multi_equals_single_launches(gpu, precision, multi, rpt, steps):
    """K iterations per launch give the iterates of K single-iteration
    launches (the same tile map and per-pixel arithmetic)."""
    shapes = [(130, 200), (64, 64), (33, 257), (1, 100), (97, 1)]
    out = {}
    for m in (1, multi):
        bat, *_ = _batch(precision, (64, rpt), m, shapes=shapes, seed=7)
        bat.solve(steps=steps, step_bound=0.124, max_iters=24, block_iters=12, tol=1e-30)
        out[m] = bat.solution()
    tol = 1e-12 if precision == "double" else 1e-6
    for a, b in zip(out[1][0] + out[1][1], out[multi][0] + out[multi][1]):
        assert np.abs(a - b).max() <= tol


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("C", [2, 3, 4])
def test_multi_channels(gpu, precision, C):
    """Per-pixel vectors of C and 2C values in the LDS region; in FP64 the
    smallest region (rows per thread 2, K = 2) exceeds the LDS budget from
    C = 3 on and is refused."""
    from opymize_amd import _hip
    shapes = [(33, 150), (9, 300), (70, 41)]
    if lds_bytes(precision, C, 2, 2) > 64 * 1024:
        with pytest.raises(_hip.HipError, match="LDS"):
            bat, *_ = _batch(precision, (64, 2), 2, shapes=shapes, C=C)
            bat.solve(steps="precond", step_bound=0.124, max_iters=4, block_iters=4, tol=1e-30)
        return
    bat, imgs, lams, data = _batch(precision, (64, 2), 2, shapes=shapes, C=C)
    k = 30
    bat.solve(steps="precond", step_bound=0.124, max_iters=k, block_iters=10, tol=1e-30)
    xs, ys = bat.solution()
    tol = 1e-9 if precision == "double" else 5e-4
    for b, im in enumerate(imgs):
        H, W = im.shape[:2]
        prob = opd.tv_problem(im.reshape(-1, C), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps="precond", step_bound=0.124, term_maxiter=k, granularity=k,
                                  term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b
        assert np.abs(ys[b] - yo).max() < tol, b


def test_multi_graph_blocks_with_odd_launch_counts(gpu):
    """Blocks of 10 iterations with K = 2 are 5 launches, so consecutive
    blocks start from alternating buffers; graph replays per start buffer
    give the eager result."""
    out = {}
    for g in (False, True):
        bat, *_ = make_batch("double")
        bat.set_multi(2)
        bat.set_tile(64, 4)
        bat.solve(steps="accel", step_bound=0.124, max_iters=70, block_iters=10, tol=1e-30, use_graph=g)
        out[g] = bat.solution()
        assert bat.cur == 0
    for a, b in zip(out[False][0] + out[False][1], out[True][0] + out[True][1]):
        np.testing.assert_array_equal(a, b)


def test_multi_converges_and_certifies(gpu):
    """A whole solve with K = 4 certifies every instance, and the strong
    convexity bound holds against the oracle's converged solution."""
    bat, imgs, lams, data = _batch("float", (64, 8), 4, seed=3, data="l2")
    res = bat.solve(steps="accel", step_bound=0.124, max_iters=20000, block_iters=40, tol=1e-5)
    xs, _ = bat.solution()
    assert all(r["status"] == 1 for r in res)
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], "l2")
        xo, _, io, _ = opd.solve(prob, steps="accel", step_bound=0.124, term_maxiter=20000, granularity=100,
                                 term_pd_gap=1e-9)
        # 0.5 lam |x - x*|^2 <= gap for either point
        bound = np.sqrt(2 * abs(res[b]["gap"]) / lams[b]) + np.sqrt(2 * abs(io["gap"]) / lams[b])
        assert np.linalg.norm(xs[b] - xo) <= bound + 1e-4, b


@pytest.mark.parametrize("multi", [1, 4])
def test_restarts_match_oracle(gpu, multi):
    """Restarts of the accelerated schedule (FP64, tests every 20
    iterations, beta 0.3, at least 40 iterations apart): the same restarts
    and iterates as the oracle's rule on the uneven batch; TV-L1 instances
    (no acceleration) never restart."""
    bat, imgs, lams, data = _batch("double", (64, 2), multi)
    k = 300
    res = bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, max_iters=k, block_iters=20, tol=1e-30,
                    restart=(0.3, 40))
    xs, ys = bat.solution()
    total = 0
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps="accel", step_bound=0.124, step_factor=4.0, term_maxiter=k,
                                  granularity=20, term_pd_gap=1e-30, restart=(0.3, 40))
        assert res[b]["restarts"] == len(io["restarts"]), (b, res[b]["restarts"], io["restarts"])
        if data[b] == "l1":
            assert res[b]["restarts"] == 0
        total += res[b]["restarts"]
        assert np.abs(xs[b] - xo).max() < 1e-9, b
        assert np.abs(ys[b] - yo).max() < 1e-9, b
    assert total > 0


def test_tv_l1_split_under_accelerated_steps(gpu):
    """step_factor_l1: under the accelerated rule the TV-L1 instances start
    from their own split (and keep constant steps), the TV-L2 ones from
    step_factor; FP64 iterates against the oracle with each instance's
    factor (its accelerated rule reduces to constant steps at gamma = 0)."""
    bat, imgs, lams, data = _batch("double", (64, 2), 4)
    k = 40
    bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, step_factor_l1=1.0 / 256, max_iters=k,
              block_iters=20, tol=1e-30, restart=False)
    xs, ys = bat.solution()
    st = bat.steps_state()
    assert "l1" in data and "l2" in data
    for b, im in enumerate(imgs):
        H, W = im.shape
        fac = 1.0 / 256 if data[b] == "l1" else 4.0
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, io, _ = opd.solve(prob, steps="accel", step_bound=0.124, step_factor=fac, term_maxiter=k,
                                  granularity=k, term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < 1e-9, b
        assert np.abs(ys[b] - yo).max() < 1e-9, b
        if data[b] == "l1":
            assert abs(st[b]["sigma"] - np.sqrt(0.124 * 256)) < 1e-12 * np.sqrt(0.124 * 256)


def test_restarts_certify_a_stalling_instance(gpu):
    """FP32, accelerated steps: the instance with lambda = 1.29 of this batch
    does not reach a certified relative gap of 1e-5 in 20000 iterations
    without restarts (its gap stops falling); with the default restarts
    every instance certifies, in fewer iterations in total."""
    out = {}
    for r in (False, "auto"):
        bat, imgs, lams, data = _batch("float", (64, 8), 4, seed=3, data="l2")
        out[r] = bat.solve(steps="accel", step_bound=0.124, max_iters=20000, block_iters=40, tol=1e-5, restart=r)
    assert out[False][0]["status"] == 0
    assert all(r["status"] == 1 for r in out["auto"])
    assert sum(r["iter"] for r in out["auto"]) < sum(r["iter"] for r in out[False])
    assert max(r["restarts"] for r in out["auto"]) > 0 and max(r["restarts"] for r in out[False]) == 0


def test_multi_rejects_bad_arguments(gpu):
    from opymize_amd import _hip
    from opymize_amd.solvers.tv import TVBatch
    bat, *_ = make_batch("float", (64, 12))
    with pytest.raises(ValueError):
        bat.set_multi(9)
    bat.multi_kernel = "tile"
    with pytest.raises(ValueError, match="streaming kernel only"):
        bat.set_multi(5)  # the tile kernel runs at most 4 iterations per launch
    with pytest.raises(ValueError):
        TVBatch([np.zeros((8, 8))], 1.0, passes=2, multi=2)
    bat.set_multi(2)  # rows per thread 12 is a single-pass tile, not a multi-iteration one (of the tile kernel)
    with pytest.raises(_hip.HipError, match="rows per thread"):
        bat.solve(steps="const", step_bound=0.124, max_iters=2, block_iters=2, tol=1e-30)
    for bad in ((1.0, 96), (0.5, -1), (0.5,)):
        with pytest.raises(ValueError, match="restart"):
            bat.solve(steps="accel", step_bound=0.124, max_iters=2, block_iters=2, tol=1e-30, restart=bad)
    with pytest.raises(ValueError, match="step_factor_l1"):
        bat.solve(steps="accel", step_bound=0.124, max_iters=2, block_iters=2, tol=1e-30, step_factor_l1=0.0)


def test_autotune_considers_iterations_per_launch(gpu):
    """With accelerated steps the tuner also times several iterations per
    launch; the winner leaves iterates that match the oracle."""
    from opymize_amd.solvers import autotune
    bat, imgs, lams, data = make_batch("double", (64, 4))
    autotune._cache.clear()
    best = autotune.tune(bat, steps="accel", iters=12, blocks=(12, 24))
    assert best["multi"] in autotune.MULTI and bat.multi == best["multi"]
    assert (bat.tw, bat.rpt) == tuple(best["tile"]) and best["block_iters"] in (12, 24)
    k = 24
    bat.solve(steps="accel", step_bound=0.124, max_iters=k, block_iters=12, tol=1e-30)
    xs, _ = bat.solution()
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, _, _, _ = opd.solve(prob, steps="accel", step_bound=0.124, term_maxiter=k, granularity=k,
                                term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < 1e-9, b
    autotune._cache.clear()


@pytest.mark.parametrize("multi,rpt", [(1, 8), (4, 4)])
def test_launch_shapes_give_the_same_iterates(gpu, multi, rpt):
    """One block per running tile (the default), the persistent grid of
    resident blocks (grid=0) and a small fixed grid (each block walks many
    tiles) give bitwise the same iterates and gaps."""
    from opymize_amd.solvers.tv import TVBatch
    imgs = [rects(H, W, 40 + i) for i, (H, W) in enumerate(SHAPES)]
    lams = [2.0, 7.5, 1.0, 15.0, 4.0, 3.3]
    out = {}
    for grid in (None, 0, 7):
        bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, rpt), grid=grid, multi=multi)
        if grid == 0:
            assert bat.grid == 0
        res = bat.solve(steps="accel", step_bound=0.124, max_iters=48, block_iters=24, tol=1e-30)
        xs, ys = bat.solution()
        out[grid] = (xs, ys, [r["gap"] for r in res])
    for g in (0, 7):
        for a, b in zip(out[None][0] + out[None][1], out[g][0] + out[g][1]):
            np.testing.assert_array_equal(a, b)
        assert out[None][2] == out[g][2]
