"""The streaming multi-iteration kernel (opy_tv_stream, TVBatch(multi_kernel=
"stream")): one wavefront per strip of 64 - 2K columns by `stream_rows` rows,
the K iterations pipelined down the rows.  Its iterates are those of the
LDS-tile kernel and of K single launches bit for bit (the same per-pixel
arithmetic with the same roundings), and those of the oracle within the
tolerances of test_gpu_tv_batch.py (FP64 1e-9, FP32 5e-4)."""
import ctypes

import numpy as np
import pytest

from oracle import pdhg as opd

from test_gpu_tv_batch import rects

pytestmark = pytest.mark.gpu

# edges of every kind: one row, one column, widths around the strip widths
# (64 - 2K = 56, 58, 60), heights around small chunk sizes
SHAPES = [(130, 200), (64, 64), (33, 257), (1, 100), (97, 1), (9, 57), (61, 59), (5, 121)]


def _batch(precision, multi, kernel, shapes=SHAPES, C=1, seed=7, data=None, rows=128, grid=None, tile=(64, 4)):
    from opymize_amd.solvers.tv import TVBatch
    rng = np.random.default_rng(seed)
    if C == 1:
        imgs = [rects(H, W, seed + i) for i, (H, W) in enumerate(shapes)]
    else:
        imgs = [np.stack([rects(H, W, 10 * C + i + k) for k in range(C)], axis=2) for i, (H, W) in enumerate(shapes)]
    lams = list(np.round(np.exp(rng.uniform(0.0, np.log(20.0), len(shapes))), 3))
    if data is None:
        data = ["l2" if i % 3 != 1 else "l1" for i in range(len(shapes))]
    bat = TVBatch(imgs, lams, data, precision=precision, tile=tile, multi=multi, multi_kernel=kernel,
                  stream_rows=rows, grid=grid)
    return bat, imgs, lams, data


def _solve(bat, steps, iters, block):
    res = bat.solve(steps=steps, step_bound=0.124, max_iters=iters, block_iters=block, tol=1e-30)
    return bat.solution(), res


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("multi", [2, 3, 4])
@pytest.mark.parametrize("steps", ["const", "accel"])
@pytest.mark.parametrize("rows", [8, 128])
def test_stream_equals_tile_kernel(gpu, precision, multi, steps, rows):
    """24 iterations in blocks of 12 on the uneven batch (TV-L1 and TV-L2
    instances): the streaming and the tile kernel give the same bits."""
    out = {}
    for kernel in ("stream", "tile"):
        bat, *_ = _batch(precision, multi, kernel, rows=rows)
        (xs, ys), res = _solve(bat, steps, 24, 12)
        out[kernel] = (xs, ys, [r["gap"] for r in res])
    for a, b in zip(out["stream"][0] + out["stream"][1], out["tile"][0] + out["tile"][1]):
        np.testing.assert_array_equal(a, b)
    assert out["stream"][2] == out["tile"][2]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("multi", [2, 4, 5, 6, 8])
def test_stream_equals_single_launches(gpu, precision, multi):
    """K iterations per streaming launch against K single-iteration launches
    (multi = 1), accelerated steps: the same bits (K = 5: two launches and two
    single iterations per block of 12)."""
    out = {}
    for m in (1, multi):
        bat, *_ = _batch(precision, m, "stream", seed=11)
        (xs, ys), _ = _solve(bat, "accel", 24, 12)
        out[m] = xs + ys
    for a, b in zip(out[1], out[multi]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("multi", [6, 8])
@pytest.mark.parametrize("rows", [8, 16, 128])
def test_stream_row_phases_equal_single_launches(gpu, precision, multi, rows):
    """Units whose columns lie inside the image run their interior steps
    unmasked and only the steps at the top and bottom image rows with row
    masks: images wide enough for such strips, from 5 rows (no unmasked
    step) to 200 (units away from both edges), give the same bits as
    single-iteration launches."""
    # widths around one and two strips (64, and 64 + 64 - 2K for K = 6, 8)
    shapes = [(200, 300), (5, 300), (21, 260), (130, 130), (17, 700), (12, 65), (30, 112), (31, 113), (9, 116),
              (11, 117)]
    out = {}
    for m in (1, multi):
        bat, *_ = _batch(precision, m, "stream", shapes=shapes, seed=13, rows=rows)
        (xs, ys), _ = _solve(bat, "accel", 48, 24)
        out[m] = xs + ys
    for a, b in zip(out[1], out[multi]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("multi", [6, 8])
def test_stream_two_channels_equal_single_launches(gpu, precision, multi):
    """Two channels (the largest C built beyond 4 iterations per launch),
    TV-L2 and TV-L1, interior strips and row phases: the same bits as
    single-iteration launches."""
    shapes = [(40, 300), (9, 200), (70, 41), (150, 160)]
    out = {}
    for m in (1, multi):
        bat, *_ = _batch(precision, m, "stream", shapes=shapes, C=2, seed=17, rows=16)
        (xs, ys), _ = _solve(bat, "accel", 48, 24)
        out[m] = xs + ys
    for a, b in zip(out[1], out[multi]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("C", [2, 3, 4])
def test_stream_channels(gpu, precision, C):
    """Per-pixel vectors of C and 2C values (Frobenius-norm TV): the streaming
    kernel against the tile kernel (bits) and the oracle."""
    shapes = [(33, 150), (9, 300), (70, 41)]
    out = {}
    for kernel in ("stream", "tile"):
        if kernel == "tile" and precision == "double" and C >= 3:
            continue  # the smallest FP64 tile region exceeds the LDS budget from C = 3 on
        bat, imgs, lams, data = _batch(precision, 2, kernel, shapes=shapes, C=C, rows=16, tile=(64, 2))
        (xs, ys), _ = _solve(bat, "accel", 30, 10)
        out[kernel] = xs + ys
    if "tile" in out:
        for a, b in zip(out["stream"], out["tile"]):
            np.testing.assert_array_equal(a, b)
    tol = 1e-9 if precision == "double" else 5e-4
    xs = out["stream"][:len(shapes)]
    for b, im in enumerate(imgs):
        H, W = im.shape[:2]
        prob = opd.tv_problem(im.reshape(-1, C), (H, W), lams[b], data[b])
        xo, _, _, _ = opd.solve(prob, steps="accel", step_bound=0.124, term_maxiter=30, granularity=30,
                                term_pd_gap=1e-30)
        assert np.abs(xs[b] - xo).max() < tol, b


def test_stream_matches_oracle_and_certifies(gpu):
    """FP64, accelerated steps with restarts: the iterates of the oracle
    after 120 iterations; then a whole FP32 solve certifies every instance in
    exactly the iterations of the tile kernel (the same iterates)."""
    bat, imgs, lams, data = _batch("double", 4, "stream", rows=32)
    k = 120
    bat.solve(steps="accel", step_bound=0.124, step_factor=4.0, max_iters=k, block_iters=20, tol=1e-30,
              restart=(0.3, 40))
    xs, ys = bat.solution()
    for b, im in enumerate(imgs):
        H, W = im.shape
        prob = opd.tv_problem(im.reshape(-1, 1), (H, W), lams[b], data[b])
        xo, yo, _, _ = opd.solve(prob, steps="accel", step_bound=0.124, step_factor=4.0, term_maxiter=k,
                                 granularity=20, term_pd_gap=1e-30, restart=(0.3, 40))
        assert np.abs(xs[b] - xo).max() < 1e-9, b
        assert np.abs(ys[b] - yo).max() < 1e-9, b
    res = {}
    for kernel in ("stream", "tile"):
        bat, *_ = _batch("float", 4, kernel, seed=3, data="l2")
        res[kernel] = bat.solve(steps="accel", step_bound=0.124, max_iters=20000, block_iters=40, tol=1e-5)
    assert all(r["status"] == 1 for r in res["stream"])
    assert [r["iter"] for r in res["stream"]] == [r["iter"] for r in res["tile"]]
    assert [r["gap"] for r in res["stream"]] == [r["gap"] for r in res["tile"]]


def test_auto_switches_to_tiles_when_few_units_run(gpu):
    """multi_kernel="auto" streams while enough units run and takes the tile
    kernel once instances drop out (here: as soon as one has certified); the
    whole solve certifies with the iterations, gaps and iterates of a solve on
    the tile kernel alone."""
    out = {}
    for kernel in ("auto", "tile"):
        bat, *_ = _batch("float", 4, kernel, seed=3, data="l2", rows=16)
        if kernel == "auto":
            bat.stream_min_units = bat.total_units  # stream only while every instance runs
            bat.solve(steps="accel", step_bound=0.124, max_iters=40, block_iters=40, tol=1e-30)
            assert bat._use_stream(bat._cfg("accel"))  # every instance still running
        bat.set_start(None, None)  # cold start for both
        res = bat.solve(steps="accel", step_bound=0.124, max_iters=20000, block_iters=40, tol=1e-5)
        out[kernel] = ([r["iter"] for r in res], [r["gap"] for r in res], bat.solution())
        if kernel == "auto":
            assert not bat._use_stream(bat._cfg("accel"))
    assert all(it < 20000 for it in out["auto"][0])
    assert out["auto"][0] == out["tile"][0] and out["auto"][1] == out["tile"][1]
    for a, b in zip(out["auto"][2][0] + out["auto"][2][1], out["tile"][2][0] + out["tile"][2][1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("multi", [4, 8])
def test_stream_units_follow_the_running_instances(gpu, multi):
    """With some instances marked certified, the compacted unit prefix covers
    exactly the running instances (host recount), the next launches are
    sized to it, and a launch leaves the stopped instances untouched (K = 8:
    the extended strip layout)."""
    import torch
    bat, imgs, lams, data = _batch("float", multi, "stream", seed=3, data="l2", rows=16)
    bat.solve(steps="accel", step_bound=0.124, max_iters=8, block_iters=8, tol=1e-30)
    stopped = [1, 4, 5]
    ck = bat.checks.view(torch.int32).reshape(bat.B, -1)
    for b in stopped:
        ck[b, 16] = 1  # status field of opy_tv_check_t: certified
    bat._compact(0)
    nrun = bat.refresh_grid()
    active = bat.active_inst.cpu().numpy()[:nrun]
    assert nrun == bat.B - len(stopped) and not set(stopped) & set(active.tolist())
    uo = bat.unit_off.cpu().numpy()
    sw = 64 - 2 * multi
    # ceil(W / sw) strips, or from K = 6 on the extended layout when it takes
    # fewer (no halo beyond the image edges: 64 + (n - 1) sw columns)
    strips = [-(-W // sw) if multi < 6 else min(-(-W // sw), 1 if W <= 64 else 1 + -(-(W - 64) // sw))
              for (H, W) in bat.shapes]
    counts = [strips[b] * (-(-bat.shapes[b][0] // 16)) for b in active]
    np.testing.assert_array_equal(uo[:nrun + 1], np.concatenate([[0], np.cumsum(counts)]))
    assert bat.sgrid * 4 >= uo[nrun] and bat.sgrid % 8 == 0
    x0 = bat.x[1 - bat.cur].clone()
    bat._multi_launch(bat.cur, bat._cfg("accel"), multi)
    torch.cuda.synchronize()
    xs = bat.unpack(bat.x[1 - bat.cur], "x")
    xo = bat.unpack(x0, "x")
    for b in stopped:  # not written by the launch
        np.testing.assert_array_equal(xs[b], xo[b])


def test_stream_graph_replays_and_launch_shapes(gpu):
    """Graph replays against eager launches, and the default grid against
    the persistent (grid=0) and a small fixed grid: the same bits."""
    out = {}
    for g, grid in ((True, None), (False, None), (True, 0), (True, 7)):
        bat, *_ = _batch("float", 4, "stream", rows=24, grid=grid)
        res = bat.solve(steps="accel", step_bound=0.124, max_iters=96, block_iters=24, tol=1e-30, use_graph=g)
        xs, ys = bat.solution()
        out[(g, grid)] = (xs + ys, [r["gap"] for r in res])
    ref = out[(True, None)]
    for key, val in out.items():
        for a, b in zip(ref[0], val[0]):
            np.testing.assert_array_equal(a, b)
        assert ref[1] == val[1], key


def test_stream_full_size_launch(gpu):
    """Two 1024 x 1024 instances (the benchmark's size), FP32, K = 4, one
    launch of each kernel from the same nontrivial iterate: the same bits on
    all 2 x 2^20 pixels (inner units at their real shape)."""
    import torch
    from opymize_amd import _hip
    print("stream full size: 2 x 1024^2, K=4, float", flush=True)
    bat, *_ = _batch("float", 4, "stream", shapes=[(1024, 1024), (1024, 1024)], data="l2", seed=2)
    bat.solve(steps="accel", step_bound=0.124, max_iters=8, block_iters=8, tol=1e-30)
    bat.refresh_grid()
    cfg = bat._cfg("accel")
    P, s = _hip.ptr, _hip.stream()
    p = bat.cur
    outs = []
    for name in ("stream", "multi"):
        if name == "stream":
            bat._call("opy_tv_stream_f32", P(bat.x[p]), P(bat.x[1 - p]), P(bat.y[p]), P(bat.y[1 - p]), P(bat.f),
                      P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.unit_off), P(bat.nact),
                      ctypes.byref(cfg), 4, bat.stream_rows, bat.sgrid, s)
        else:
            bat._call("opy_tv_multi_f32", P(bat.x[p]), P(bat.x[1 - p]), P(bat.y[p]), P(bat.y[1 - p]), P(bat.f),
                      P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.act_off), P(bat.nact),
                      ctypes.byref(cfg), 4, bat.grid, s)
        torch.cuda.synchronize()
        outs.append((bat.x[1 - p].clone(), bat.y[1 - p].clone()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert float(outs[0][0].abs().max()) > 0


@pytest.mark.parametrize("multi", [6, 8])
def test_stream_many_iterations_match_oracle(gpu, multi):
    """Beyond the tile kernel's 4 iterations per launch: FP64 iterates of the
    oracle after 96 iterations (accelerated, restarts), C = 1 and C = 2."""
    for C, shapes in ((1, SHAPES), (2, [(33, 150), (9, 300), (70, 41)])):
        bat, imgs, lams, data = _batch("double", multi, "stream", shapes=shapes, C=C, rows=24)
        k = 96
        bat.solve(steps="accel", step_bound=0.124, max_iters=k, block_iters=48, tol=1e-30, restart=(0.3, 48))
        xs, ys = bat.solution()
        for b, im in enumerate(imgs):
            H, W = im.shape[:2]
            prob = opd.tv_problem(im.reshape(-1, C), (H, W), lams[b], data[b])
            xo, yo, _, _ = opd.solve(prob, steps="accel", step_bound=0.124, term_maxiter=k, granularity=48,
                                     term_pd_gap=1e-30, restart=(0.3, 48))
            assert np.abs(xs[b] - xo).max() < 1e-9, (C, b)
            assert np.abs(ys[b] - yo).max() < 1e-9, (C, b)


def test_stream_rejects_bad_arguments(gpu):
    from opymize_amd import _hip
    from opymize_amd.solvers.tv import TVBatch
    with pytest.raises(ValueError, match="streaming kernel only"):
        TVBatch([np.zeros((8, 8))], 1.0, multi=6, multi_kernel="tile")
    with pytest.raises(ValueError, match="streaming kernel only"):
        TVBatch([np.zeros((8, 8, 3))], 1.0, multi=6)
    with pytest.raises(ValueError):
        TVBatch([np.zeros((8, 8))], 1.0, multi=9)
    six = TVBatch([np.zeros((8, 8))], 1.0, multi=6)
    with pytest.raises(ValueError, match="preconditioned"):
        six.solve(steps="precond", step_bound=0.124, max_iters=12, block_iters=12, tol=1e-30)
    with pytest.raises(ValueError):
        TVBatch([np.zeros((8, 8))], 1.0, multi=2, multi_kernel="lds")
    with pytest.raises(ValueError):
        TVBatch([np.zeros((8, 8))], 1.0, multi=2, stream_rows=0)
    # an instance whose y block would pass 4 GiB (32-bit byte offsets): tiles under "auto",
    # refused when streaming is forced (the shapes are stood in, nothing is launched)
    big = TVBatch([np.zeros((8, 8))], 1.0, multi=4)
    cfg = big._cfg("accel")
    big._stream_now = True
    assert big._use_stream(cfg)
    big.shapes = [(23171, 23171)]  # 2 x 4 B x 23171^2 > 2^32
    assert not big._stream_fits() and not big._use_stream(cfg)
    big.multi_kernel = "stream"
    with pytest.raises(ValueError, match="32-bit"):
        big._use_stream(cfg)
    bat, *_ = _batch("float", 4, "stream")
    with pytest.raises(ValueError, match="preconditioned"):
        bat.solve(steps="precond", step_bound=0.124, max_iters=8, block_iters=8, tol=1e-30)
    bat.solve(steps="accel", step_bound=0.124, max_iters=4, block_iters=4, tol=1e-30)
    P, s = _hip.ptr, _hip.stream()

    def call(cfg, k, rows, sfx="_f32"):
        bat._call("opy_tv_stream" + sfx, P(bat.x[0]), P(bat.x[1]), P(bat.y[0]), P(bat.y[1]), P(bat.f),
                  P(bat.insts), P(bat.state), P(bat.active_inst), P(bat.unit_off), P(bat.nact), ctypes.byref(cfg),
                  k, rows, bat.sgrid, s)

    with pytest.raises(_hip.HipError, match="preconditioned"):
        call(bat._cfg("precond"), 4, 128)
    with pytest.raises(_hip.HipError, match="k = 9"):
        call(bat._cfg("accel"), 9, 128)
    c3 = bat._cfg("accel")
    c3.C = 3
    with pytest.raises(_hip.HipError, match="C <= 2"):
        call(c3, 6, 128)
    with pytest.raises(_hip.HipError, match="rows"):
        call(bat._cfg("accel"), 4, 0)
    with pytest.raises(_hip.HipError, match="rows"):
        bat._call("opy_tv_stream_units", P(bat.insts), P(bat.active_inst), P(bat.nact), P(bat.unit_off), 4, -1, s)
    # the auto choice keeps preconditioned steps on the tile kernel (whatever the unit count)
    auto, *_ = _batch("float", 4, "auto")
    auto.stream_min_units = 0
    auto.solve(steps="precond", step_bound=0.124, max_iters=8, block_iters=8, tol=1e-30)
    assert not auto._use_stream(auto._cfg("precond")) and auto._use_stream(auto._cfg("accel"))
