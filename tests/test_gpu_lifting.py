"""The multi-label lifting workload on the GPU (opymize_amd.lifting, solved
by the generic PDHG path with preconditioned steps) against the oracle
(oracle/lifting.py).

Tolerances: FP64 iterates to 1e-9 of the FP64 oracle after a fixed number of
iterations (summation-order differences only); FP32 iterates to 5e-4
absolute (per-iteration rounding of FP32 arithmetic, including the matrix
cores' FP32 products); the gap to 1e-9 relative in FP64."""
import numpy as np
import pytest

from oracle import lifting as olf
from oracle import pdhg as opd
from opymize_amd.tools.synth import rects

pytestmark = pytest.mark.gpu


def _solve_gpu(f, L, lam, precision, iters, gran=None, use_graph=True, fused=False):
    from opymize_amd import lifting
    from opymize_amd.solvers import PDHG
    G, F, A, gam = lifting.sublabel_tv(f, L, lam)
    s = PDHG(G, F, A)
    info = s.solve(steps="precond", precision=precision, term_maxiter=iters, granularity=gran or iters,
                   term_pd_gap=1e-30, use_graph=use_graph, fused=fused)
    x, y = s.state
    return np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), info, gam


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("L", [2, 16])
def test_lifting_iterates_match_oracle(gpu, precision, L, fused):
    """The generic chain (per-operator kernels) and the fused batched kernels."""
    f = rects(20, 17, 1, seed=11).reshape(20, 17)
    lam, iters = 0.25, 120
    x, y, info, gam = _solve_gpu(f, L, lam, precision, iters, fused=fused)
    prob, ogam = olf.problem(f, L, lam)
    assert np.allclose(gam, ogam, rtol=0, atol=1e-15)
    xo, yo, io, _ = opd.solve(prob, steps="precond", term_maxiter=iters, granularity=iters, term_pd_gap=1e-30)
    tol = 1e-9 if precision == "double" else 5e-4
    assert np.abs(x - xo).max() < tol
    assert np.abs(y - yo).max() < tol * max(1.0, np.abs(yo).max())
    if precision == "double":
        assert abs(info["gap"] - io["gap"]) < 1e-9 * max(1.0, abs(io["objp"]))
        assert abs(info["objp"] - io["objp"]) < 1e-9 * max(1.0, abs(io["objp"]))


def test_lifting_full_size_first_iterations(gpu):
    """The 512 x 512, L = 16 workload in FP32 for 6 iterations against the
    FP64 oracle at full size (every kernel of the iteration at its real
    shape: the (512, 512, 15) gradient, the 15 x 30 per-pixel products on the
    matrix cores, 3.9 M epigraph projections, 262144 simplex projections)."""
    print("lifting full size: 512x512 L=16 float, 6 iterations", flush=True)
    f = rects(512, 512, 1, seed=4).reshape(512, 512)
    lam, iters = 0.25, 6
    x, y, info, gam = _solve_gpu(f, 16, lam, "float", iters)
    prob, _ = olf.problem(f, 16, lam)
    xo, yo, io, _ = opd.solve(prob, steps="precond", term_maxiter=iters, granularity=iters, term_pd_gap=1e-30)
    assert np.abs(x - xo).max() < 5e-4
    assert np.abs(y - yo).max() < 5e-4 * max(1.0, np.abs(yo).max())
    assert abs(info["objp"] - io["objp"]) < 1e-4 * abs(io["objp"])


def test_lifting_graph_replays_match_eager(gpu):
    f = rects(16, 12, 1, seed=3).reshape(16, 12)
    a = _solve_gpu(f, 8, 0.25, "double", 40, use_graph=True)
    b = _solve_gpu(f, 8, 0.25, "double", 40, use_graph=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_lifting_solution_matches_rof(gpu):
    """Convexity check: for the convex data (s - f)^2 the unlifted s of the
    lifted solution is the ROF solution within the certificate
    |s - s_rof| <= sqrt(P_lift - D_rof) + sqrt(gap_rof) (the energy is
    2-strongly convex, bounded above by the lifted primal value and below by
    the ROF dual value), and the lifted problem is solved to a small gap."""
    from opymize_amd import lifting
    f = rects(24, 24, 1, seed=8).reshape(24, 24)
    lam = 0.25
    x, y, info, gam = _solve_gpu(f, 16, lam, "double", 40000, gran=4000)
    s = lifting.unlift(x, gam, 24 * 24)
    rp = opd.tv_problem(f.reshape(-1, 1), (24, 24), 2.0, tv_lbd=lam)
    xr, yr, ir, _ = opd.solve(rp, steps="adaptive", step_bound=0.124, term_maxiter=20000, granularity=20000,
                              term_pd_gap=1e-30)
    bound = np.sqrt(max(info["objp"] - ir["objd"], 0.0)) + np.sqrt(max(ir["gap"], 0.0))
    err = np.linalg.norm(s - xr)
    print("lifting vs ROF: relgap %.2e, |s - s_rof| = %.3e <= %.3e" % (info["relgap"], err, bound))
    assert abs(info["relgap"]) < 1e-4
    assert err <= bound
    assert bound < 0.2 * np.linalg.norm(xr - f.ravel())


def test_lifting_batch_uneven_sizes_and_labels(gpu):
    """One batch of instances of uneven image size and label count (2 to 33
    labels, tiles cut by the image edges) against per-instance oracle runs
    after a fixed number of iterations, in FP64; every instance's gap as the
    oracle's."""
    from opymize_amd.solvers.lift import LiftBatch
    shapes = [(20, 17), (9, 70), (33, 12), (5, 130), (64, 65)]
    Ls = [16, 5, 2, 33, 9]
    lams = [0.25, 0.1, 0.4, 0.2, 0.3]
    imgs = [rects(H, W, 1, seed=20 + i).reshape(H, W) for i, (H, W) in enumerate(shapes)]
    bat = LiftBatch(imgs, Ls, lams, precision="double")
    iters = 60
    res = bat.solve(max_iters=iters, block_iters=20, check_every=3, tol=1e-30)
    xs, ys = bat.solution()
    for b, im in enumerate(imgs):
        prob, gam = olf.problem(im, Ls[b], lams[b])
        xo, yo, io, _ = opd.solve(prob, steps="precond", term_maxiter=iters, granularity=iters, term_pd_gap=1e-30)
        print("instance %d: %s L=%d max |x - xo| %.2e" % (b, shapes[b], Ls[b], np.abs(xs[b] - xo).max()))
        assert np.abs(xs[b] - xo).max() < 1e-9, b
        assert np.abs(ys[b] - yo).max() < 1e-9 * max(1.0, np.abs(yo).max()), b
        assert res[b]["iter"] == iters
        assert abs(res[b]["gap"] - io["gap"]) < 1e-9 * max(1.0, abs(io["objp"])), (b, res[b]["gap"], io["gap"])
        assert abs(res[b]["objd"] - io["objd"]) < 1e-9 * max(1.0, abs(io["objp"])), b


def test_lifting_batch_stops_certified_instances(gpu):
    """Instances that certify stop iterating while the others go on; each
    stops at a check with its gap below the tolerance."""
    from opymize_amd.solvers.lift import LiftBatch
    imgs = [rects(12, 10, 1, seed=5).reshape(12, 10), rects(24, 30, 1, seed=6).reshape(24, 30)]
    bat = LiftBatch(imgs, [2, 8], [0.25, 0.25], precision="double")
    res = bat.solve(max_iters=20000, block_iters=100, check_every=1, tol=1e-8)
    assert res[0]["status"] == 1 and abs(res[0]["relgap"]) < 1e-8
    assert res[0]["iter"] < res[1]["iter"] or res[1]["status"] != 1


@pytest.mark.parametrize("fused", [False, True])
def test_lifting_primal_weight(gpu, fused):
    """precond_weight w scales the primal steps by w and the dual ones by 1/w,
    in the generic chain and the fused kernels alike (FP64 against the oracle)."""
    from opymize_amd import lifting
    from opymize_amd.solvers import PDHG
    f = rects(18, 21, 1, seed=13).reshape(18, 21)
    G, F, A, gam = lifting.sublabel_tv(f, 12, 0.25)
    s = PDHG(G, F, A)
    s.solve(steps="precond", precision="double", term_maxiter=80, granularity=80, term_pd_gap=1e-30, fused=fused,
            precond_weight=6.0)
    x, y = s.state
    prob, _ = olf.problem(f, 12, 0.25)
    xo, yo, io, _ = opd.solve(prob, steps="precond", term_maxiter=80, granularity=80, term_pd_gap=1e-30,
                              precond_weight=6.0)
    xo1, _, _, _ = opd.solve(prob, steps="precond", term_maxiter=80, granularity=80, term_pd_gap=1e-30)
    assert np.abs(x - xo).max() < 1e-9 and np.abs(y - yo).max() < 1e-9 * max(1.0, np.abs(yo).max())
    assert np.abs(xo - xo1).max() > 1e-3  # the weight changes the iterates
