"""Pins of the oracle's multi-label lifting model (oracle/lifting.py), none of
which retypes the model's formulas: the simplex projection against its KKT
conditions and an independent QP solve, exactness of the lifted data term
at lifted points, the lifted energy as an upper bound of the unlifted one,
the one-piece model against the ROF oracle, and the block-constant
preconditioner against the reference's per-element one."""
import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import lifting as lf
from oracle import pdhg as opd
from opymize_amd.tools.synth import rects


def test_simplex_projection_kkt_and_qp():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 7, 15):
        V = rng.normal(0, 1.5, (40, n))
        P = lf.simplex_proj(V)
        assert np.all(P >= 0) and np.allclose(P.sum(axis=1), 1.0, atol=1e-12)
        # KKT: V - P = theta on the support, V - P <= theta off it
        for v, p in zip(V, P):
            sup = p > 0
            theta = (v - p)[sup]
            assert np.ptp(theta) < 1e-12
            assert np.all(v[~sup] <= theta[0] + 1e-12)
        for v, p in zip(V[:5], P[:5]):
            res = minimize(lambda q: 0.5 * ((q - v) ** 2).sum(), np.full(n, 1.0 / n), jac=lambda q: q - v,
                           bounds=[(0, None)] * n, constraints=[{"type": "eq", "fun": lambda q: q.sum() - 1}],
                           method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
            assert np.abs(res.x - p).max() < 1e-6
    # points already in the simplex are fixed; a constant shift does not matter
    q = rng.dirichlet(np.ones(6), 10)
    assert np.allclose(lf.simplex_proj(q), q, atol=1e-14)
    assert np.allclose(lf.simplex_proj(q + 3.7), q, atol=1e-12)


@pytest.mark.parametrize("L", [2, 5, 16])
def test_data_term_exact_at_lifted_points(L):
    """At a lifted point (U, r = e_i) the perspective pieces sum to (s - f)^2."""
    rng = np.random.default_rng(L)
    H, W = 6, 7
    f = rng.uniform(0, 1, (H, W))
    lam = 0.3
    prob, gam = lf.problem(f, L, lam, lo=-0.2, hi=1.3)
    s = rng.uniform(gam[0], gam[-1], H * W)
    U, r = lf.lift(s, gam)
    assert np.allclose(lf.unlift(U, gam), s, atol=1e-13)
    N, k = H * W, L - 1
    Kx = prob.A.dot(np.concatenate([U.ravel(), r.ravel()]))
    z = Kx[3 * N * k:].reshape(N * k, 2)
    # only the piece that holds s has r_i = 1; its pair is (s, -1)
    zi = z.reshape(N, k, 2)[np.arange(N), np.argmax(r, axis=1)]
    assert np.allclose(zi[:, 0], s, atol=1e-12) and np.all(zi[:, 1] == -1)
    # the other pairs are exactly zero, so they add nothing
    mask = np.ones((N, k), bool)
    mask[np.arange(N), np.argmax(r, axis=1)] = False
    assert not mask.any() or np.abs(z.reshape(N, k, 2)[mask]).max() < 1e-12
    val, inf = prob.F(np.concatenate([np.zeros(3 * N * k), z.ravel()]))
    assert abs(val - ((s - f.ravel()) ** 2).sum()) < 1e-10


def test_lifted_energy_bounds_unlifted():
    """For every feasible (U, r) with r > 0 the lifted primal value is at
    least sum (s - f)^2 + lam TV(s) of s = gam_1 + h sum U; with one piece it
    is equal."""
    rng = np.random.default_rng(3)
    H, W = 5, 6
    f = rng.uniform(0, 1, (H, W))
    for L in (2, 4, 9):
        prob, gam = lf.problem(f, L, 0.4, lo=0.0, hi=1.0)
        N, k = H * W, L - 1
        for _ in range(20):
            U = rng.uniform(0, 1, (N, k))
            r = rng.dirichlet(np.ones(k), N)
            x = np.concatenate([U.ravel(), r.ravel()])
            P, inf = prob.F(prob.A.dot(x))
            E = lf.unlifted_energy(lf.unlift(U, gam), f, 0.4)
            assert inf == 0 and P >= E - 1e-10
            if L == 2:
                assert abs(P - E) < 1e-10


def _rof(f, lam, iters=20000):
    H, W = f.shape
    rp = opd.tv_problem(f.reshape(-1, 1), (H, W), 2.0, tv_lbd=lam)  # sum (s - f)^2 = 0.5*2*|s - f|^2
    xs, ys, info, _ = opd.solve(rp, steps="adaptive", step_bound=0.124, term_maxiter=iters, granularity=iters,
                                term_pd_gap=1e-30)
    return xs, info


def test_one_piece_is_rof():
    """With L = 2 the lifted problem is ROF restricted to [gam_1, gam_2]; with
    labels spanning f's range the restriction is inactive."""
    f = rects(12, 10, 1, seed=5).reshape(12, 10)
    lam = 0.25
    prob, gam = lf.problem(f, 2, lam)
    x, y, info, _ = opd.solve(prob, steps="precond", term_maxiter=20000, granularity=20000, term_pd_gap=1e-30)
    assert abs(info["relgap"]) < 1e-10
    s = lf.unlift(x[:120], gam)
    xs, ir = _rof(f, lam)
    assert abs(ir["relgap"]) < 1e-10
    assert np.abs(s - xs).max() < 1e-6
    assert abs(info["objp"] - lf.unlifted_energy(xs, f, lam)) < 1e-8


def test_block_constant_steps_needed():
    """The reference's per-element diagonal steps with the Euclidean epigraph
    and simplex projections stall at a point that is not a solution; the
    block-constant steps converge (one-piece model, where the solution is
    the ROF one)."""
    f = rects(10, 10, 1, seed=7).reshape(10, 10)
    prob, gam = lf.problem(f, 2, 0.2)
    xs, ir = _rof(f, 0.2)
    target = lf.unlifted_energy(xs, f, 0.2)
    _, _, good, _ = opd.solve(prob, steps="precond", term_maxiter=4000, granularity=4000, term_pd_gap=1e-30)
    prob.xblocks, prob.yblocks = [], []
    _, _, bad, _ = opd.solve(prob, steps="precond", term_maxiter=4000, granularity=4000, term_pd_gap=1e-30)
    assert abs(good["objp"] - target) < 1e-6 and abs(good["relgap"]) < 1e-6
    assert bad["relgap"] > 1e-2 and bad["objp"] - target > 1e-2


def test_lifted_solution_within_rof_certificate():
    """s of a lifted iterate x satisfies |s - s_opt|^2 <= E(s) - E(s_opt)
    <= P_lift(x) - D_ROF (E is 2-strongly convex, E(s) <= P_lift(x), and the
    ROF dual value D_ROF <= E(s_opt)); checked with L = 6 after enough
    iterations that this bound is small."""
    f = rects(10, 9, 1, seed=9).reshape(10, 9)
    lam = 0.3
    prob, gam = lf.problem(f, 6, lam)
    x, y, info, _ = opd.solve(prob, steps="precond", term_maxiter=30000, granularity=30000, term_pd_gap=1e-30)
    s = lf.unlift(x[:90 * 5], gam)
    xs, ir = _rof(f, lam)
    bound = np.sqrt(max(info["objp"] - ir["objd"], 0.0)) + np.sqrt(max(ir["gap"], 0.0))
    assert np.linalg.norm(s - xs) <= bound + 1e-9
    assert bound < 0.5  # the relaxation is nearly tight for this image
    # weak duality of the lifted problem and the ordering of the optima
    assert info["objd"] <= info["objp"] + 1e-12
    assert ir["objd"] <= info["objp"] + 1e-9
