"""Pins of the oracle PDHG (CPU only)."""
import numpy as np
import pytest

from oracle import pdhg, tv1d


def test_known_answer_three_iterations():
    # 1x2 image, f = (0, 1), SSD volume 1, TV weight 1, constant steps with
    # bound 1/4 -> sigma = tau = 1/2.  Worked by hand from the PDHG update
    # x+ = prox(x - tau A^T y), y+ = proj(y + sigma (2 A x+ - A x)):
    # it1: x = (0, 1/3), y = 1/3; it2: x = (1/9, 4/9), y = 1/2; it3: x = (13/54, 25/54), y = 5/9
    f = np.array([[0.0], [1.0]])
    prob = pdhg.tv_problem(f, (1, 2), 1.0)
    expect = {1: ([0, 1 / 3], 1 / 3), 2: ([1 / 9, 4 / 9], 1 / 2), 3: ([13 / 54, 25 / 54], 5 / 9)}
    for k, (xe, ye) in expect.items():
        x, y, info, _ = pdhg.solve(prob, steps="const", step_bound=0.25, term_maxiter=k, granularity=10 ** 9)
        np.testing.assert_allclose(x, xe, atol=1e-15)
        np.testing.assert_allclose(y, [0, ye, 0, 0], atol=1e-15)


def _rects(H, W, seed):
    rng = np.random.default_rng(seed)
    img = np.full((H, W), 0.5)
    for _ in range(5):
        y0, y1 = sorted(rng.integers(0, H, 2))
        x0, x1 = sorted(rng.integers(0, W, 2))
        img[y0:y1 + 1, x0:x1 + 1] = rng.random()
    return img + 0.1 * rng.standard_normal((H, W))


@pytest.mark.parametrize("steps", ["const", "adaptive", "precond", "accel"])
def test_1d_tv_converges_to_exact_solution(steps):
    W, lam = 40, 4.0
    f = _rects(1, W, 0).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (1, W), lam)
    x, y, info, _ = pdhg.solve(prob, steps=steps, step_bound=0.124, term_maxiter=20000, granularity=100,
                               term_pd_gap=1e-12)
    xs, _ = tv1d.tv1d_denoise(f.ravel(), 1.0 / lam)
    # strong convexity of G: 0.5*lam*|x - x*|^2 <= gap
    assert info["gap"] >= -1e-10
    assert np.linalg.norm(x - xs) <= np.sqrt(2 * max(info["gap"], 0) / lam) + 1e-6


def test_weak_duality_and_fixed_point_2d():
    H, W, lam = 12, 14, 6.0
    f = _rects(H, W, 1).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (H, W), lam)
    gaps = []
    for k in (10, 100, 400):
        x, y, info, _ = pdhg.solve(prob, steps="const", step_bound=0.124, term_maxiter=k, granularity=k)
        assert info["gap"] >= -1e-12 and info["infeasd"] <= 1e-12
        gaps.append(info["gap"])
    assert gaps[2] < gaps[1] < gaps[0] and gaps[-1] < 0.05 * gaps[0]
    # fixed point of the PDHG map at convergence
    x, y, info, _ = pdhg.solve(prob, steps="const", step_bound=0.124, term_maxiter=6000, granularity=6000)
    tau = sigma = np.sqrt(0.124)
    xn = prob.gprox(x - tau * prob.AT.dot(y), tau)
    yn = prob.fconjprox(y + sigma * prob.A.dot(2 * xn - x), sigma)
    assert np.abs(xn - x).max() < 1e-4 and np.abs(yn - y).max() < 1e-4


def test_adaptive_rule_invariants():
    H, W, lam = 10, 9, 3.0
    f = _rects(H, W, 2).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (H, W), lam)
    bnd = 0.124
    x, y, info, hist = pdhg.solve(prob, steps="adaptive", step_bound=bnd, term_maxiter=300, granularity=10 ** 9,
                                  history=True)
    taus = np.array([h[0] for h in hist])
    sigmas = np.array([h[1] for h in hist])
    alphas = np.array([h[2] for h in hist])
    # tau*sigma stays at the bound; alpha only shrinks by eta; steps change by (1-alpha)^{+-1}
    np.testing.assert_allclose(taus * sigmas, bnd, rtol=1e-12)
    ch = np.flatnonzero(np.diff(alphas) != 0)
    np.testing.assert_allclose(alphas[ch + 1] / alphas[ch], 0.95, rtol=1e-12)
    r = taus[ch + 1] / taus[ch]
    ok = np.isclose(r, 1 / (1 - alphas[ch])) | np.isclose(r, 1 - alphas[ch])
    assert ok.all() and len(ch) > 0


def test_l1_problem_gap_certifies():
    H, W, lam = 8, 8, 1.5
    f = _rects(H, W, 3).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (H, W), lam, data="l1")
    x, y, info, _ = pdhg.solve(prob, steps="const", step_bound=0.124, term_maxiter=3000, granularity=3000)
    assert info["gap"] >= -1e-9 or info["infeasp"] > 0


def test_truncate_examples():
    # reference tools/__init__.py:6 examples
    assert pdhg.truncate(0.006142, 3) == 0.00614
    assert pdhg.truncate(2341.2, 3) == 2340.0
    from opymize_amd.tools import truncate
    for v in [0.006142, 2341.2, 0.12499, 0.5, 7.77777, 123456.0, -0.0412345]:
        assert truncate(v, 3) == pdhg.truncate(v, 3)


def test_accelerated_steps_rate_and_reduction():
    """Accelerated steps (Chambolle & Pock 2011, Algorithm 2): the squared
    primal error against the exact 1-D TV solution is O(1/N^2) (their
    Theorem 2), so N * |x_N - x*| stays bounded; with gamma = 0 the iteration
    is the constant-step one.  (On this small problem the constant and
    adaptive steps converge linearly and end up ahead; the accelerated rule
    pays off on large images within the iteration budgets used here.)"""
    W, lam = 60, 2.0
    f = _rects(1, W, 3).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (1, W), lam)
    xs, _ = tv1d.tv1d_denoise(f.ravel(), 1.0 / lam)
    err = {}
    for steps in ("accel",):
        err[steps] = []
        for n in (100, 200, 400, 800):
            x, _, _, _ = pdhg.solve(prob, steps=steps, step_bound=0.124, term_maxiter=n, granularity=10 ** 9)
            err[steps].append(np.linalg.norm(x - xs))
    e = np.array(err["accel"])
    # N * error stays bounded and the error keeps falling
    assert e[-1] * 800 <= 1.2 * e[0] * 100
    assert np.all(np.diff(e) < 0)
    prob.gamma = 0.0
    xa, ya, _, _ = pdhg.solve(prob, steps="accel", step_bound=0.124, term_maxiter=50, granularity=10 ** 9)
    xc, yc, _, _ = pdhg.solve(prob, steps="const", step_bound=0.124, term_maxiter=50, granularity=10 ** 9)
    np.testing.assert_allclose(xa, xc, rtol=0, atol=1e-14)
    np.testing.assert_allclose(ya, yc, rtol=0, atol=1e-14)


def test_restart_equals_fresh_solve_from_the_iterate():
    """A restart of the accelerated schedule keeps (x, y) and returns tau,
    sigma to their initial values: the restarted run equals a plain run up
    to the restart followed by a fresh plain run started from that iterate,
    and restarts fall on tests at least min_iters apart."""
    H, W, lam = 12, 14, 1.5
    f = _rects(H, W, 5).reshape(-1, 1)
    prob = pdhg.tv_problem(f, (H, W), lam)
    kw = dict(steps="accel", step_bound=0.124, step_factor=4.0, term_pd_gap=1e-30, granularity=10)
    x, y, info, _ = pdhg.solve(prob, term_maxiter=400, restart=(0.3, 20), **kw)
    rs = info["restarts"]
    assert len(rs) >= 2, rs
    assert all(r % 10 == 0 for r in rs) and rs[0] >= 20 and np.all(np.diff(rs) >= 20), rs
    k1, k2 = rs[0], rs[1]
    x1, y1, i1, _ = pdhg.solve(prob, term_maxiter=k1, **kw)
    x2, y2, i2, _ = pdhg.solve(prob, term_maxiter=k2 - k1, x0=x1, y0=y1, **kw)
    xr, yr, ir, _ = pdhg.solve(prob, term_maxiter=k2, restart=(0.3, 20), **kw)
    assert ir["restarts"] == [k1, k2]
    np.testing.assert_array_equal(xr, x2)
    np.testing.assert_array_equal(yr, y2)
    # the plain run diverges from the restarted one after the first restart
    xp, _, _, _ = pdhg.solve(prob, term_maxiter=k2, **kw)
    assert np.abs(xp - xr).max() > 1e-8


def test_restart_rule_on_a_gap_sequence():
    """The firing rule against a hand-worked sequence of tests: with
    min_iters = 20 and tests every 10 iterations, the first test records the
    gap only, a rise above the smallest gap since the last restart or a fall
    below beta times the gap at it fires, and not within 20 iterations of
    the last restart."""
    def rule(gaps, beta=0.3, min_iters=20, every=10):
        ref, best, last, out = np.inf, np.inf, 0, []
        for i, g in enumerate(gaps):
            it = (i + 1) * every
            fire = it - last >= min_iters and ((np.isfinite(ref) and g < beta * ref) or g > best)
            best = min(best, g)
            if not np.isfinite(ref):
                ref = g
            if fire:
                ref, best, last = g, np.inf, it
                out.append(it)
        return out
    # 10: record ref 1.0; 20: 0.5 no (0.5 >= 0.3, no rise); 30: 0.6 > 0.5 rise -> fire;
    # 40: within 20 of 30; 50: 0.1 < 0.3 * 0.6 -> fire; 60, 70: falling slowly; 80: rise -> fire
    gaps = [1.0, 0.5, 0.6, 0.55, 0.1, 0.09, 0.085, 0.2]
    assert rule(gaps) == [30, 50, 80]
    # the oracle's solve applies the same rule: replay its own test values
    H, W, lam = 10, 9, 1.2
    prob = pdhg.tv_problem(_rects(H, W, 2).reshape(-1, 1), (H, W), lam)
    seen = []
    orig = pdhg.pd_gap

    def spy(*a):
        r = orig(*a)
        seen.append(abs(r["relgap"]))
        return r
    pdhg.pd_gap = spy
    try:
        _, _, info, _ = pdhg.solve(prob, steps="accel", step_bound=0.124, step_factor=4.0, term_pd_gap=1e-30,
                                   term_maxiter=300, granularity=10, restart=(0.3, 20))
    finally:
        pdhg.pd_gap = orig
    assert info["restarts"] and info["restarts"] == rule(seen)
