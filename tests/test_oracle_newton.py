"""Pins of the semismooth Newton oracle (oracle/newton.py) that do not
restate its formulas: central differences of the prox maps and of the
fixed-point residual, a direct solve, a hand-derived line search and the
exact 1-D TV solution."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import newton, pdhg, prox, tv1d


def _fd_jacobian(fun, x, h=1e-6):
    """Central-difference Jacobian of fun at x (dense)."""
    x = np.asarray(x, dtype=np.float64)
    f0 = fun(x)
    J = np.zeros((f0.size, x.size))
    for j in range(x.size):
        e = np.zeros_like(x)
        e[j] = h
        J[:, j] = (fun(x + e) - fun(x - e)) / (2 * h)
    return J


def _away_from(x, edges, margin):
    return np.all([np.all(np.abs(x - e) > margin) for e in edges])


def test_l12proj_jacobian_matches_differences():
    rng = np.random.default_rng(0)
    N, M, lbd = 30, 4, 1.0
    xb = rng.standard_normal((N, M)) * 0.8
    n = np.sqrt((xb ** 2).sum(axis=1))
    xb[np.abs(n - lbd) < 0.05] *= 1.3  # keep every point off the sphere
    n = np.sqrt((xb ** 2).sum(axis=1))
    assert (n > lbd).any() and (n < lbd).any()
    J = newton.l12proj_jacobian(xb.ravel(), N, M, lbd).toarray()
    Jfd = _fd_jacobian(lambda v: prox.l1norms_proj(v, N, M, 1, lbd), xb.ravel())
    np.testing.assert_allclose(J, Jfd, atol=1e-8)
    # symmetric, identity inside, a contraction with kernel u outside
    np.testing.assert_allclose(J, J.T, atol=1e-14)
    i = int(np.argmax(n))
    Bi = J[i * M:(i + 1) * M, i * M:(i + 1) * M]
    np.testing.assert_allclose(Bi @ xb[i], 0.0, atol=1e-14)


def test_signproj_and_ssd_jacobians_match_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(40)
    x[np.abs(x) < 0.01] = 0.5
    for positive, fun in ((True, lambda v: np.maximum(v, 0)), (False, lambda v: np.minimum(v, 0))):
        J = newton.signproj_jacobian(x, positive).toarray()
        np.testing.assert_allclose(J, _fd_jacobian(fun, x), atol=1e-8)
    N, C = 10, 3
    f = rng.standard_normal((N, C))
    vol = np.array([0.5, 2.0, 3.0])
    mask = rng.random(N) < 0.7
    for tau in (0.3, rng.random(N * C) + 0.1):
        J = newton.ssd_prox_jacobian(N, C, vol, tau, mask).toarray()
        Jfd = _fd_jacobian(lambda v: prox.ssd_prox(v, f, vol, tau, mask), rng.standard_normal(N * C))
        np.testing.assert_allclose(J, Jfd, atol=1e-8)


def test_newton_matrix_is_the_residual_jacobian():
    """M = [[I - K, tau K A^T], [-sigma H A, I - H]] against central
    differences of F(x, y) = (x - prox(x - tau A^T y), y - prox*(y + sigma A x))."""
    rng = np.random.default_rng(2)
    dims, C = (5, 4), 2
    N = 20
    f = rng.random((N, C))
    prob = newton.tv_newton_problem(f, dims, lam=3.0, tv_lbd=0.2)
    tau, sigma = 0.3, 0.35
    xy = np.concatenate([rng.random(prob.n), 0.15 * rng.standard_normal(prob.m)])
    F, K, H = newton.residual(prob, xy, tau, sigma)
    M = newton.newton_matrix(prob.A, K, H, tau, sigma).toarray()
    Mfd = _fd_jacobian(lambda v: newton.residual(prob, v, tau, sigma)[0], xy)
    np.testing.assert_allclose(M, Mfd, atol=1e-7)


def test_cg_matches_direct_solve():
    rng = np.random.default_rng(3)
    n = 40
    Mmat = sp.csr_matrix(rng.standard_normal((n, n)) / np.sqrt(n) + np.eye(n))
    b = rng.standard_normal(n)
    for lbd in (0.0, 0.5):
        x, steps = newton.cg_solve(Mmat, b, np.zeros(n), lbd=lbd, maxiter=1000, tol=1e-26)
        A = Mmat.toarray()
        xs = np.linalg.solve(A.T @ A + lbd * np.eye(n), -A.T @ b)
        np.testing.assert_allclose(x, xs, rtol=1e-9, atol=1e-10)
        assert steps <= 1001


def test_armijo_backtracking_by_hand():
    """f(x) = x^2/2 from x = 1 along p = -4 with slope -4: alpha = 1, 1/2,
    1/4, 1/8, 1/16 give f = 4.5, 0.5, 0, 0.125, 0.28125, above 0.5 - 3.6 alpha;
    alpha = 1/32 gives x = 0.875, f = 0.3828125 <= 0.3875."""
    fun = lambda v: 0.5 * float(v @ v)
    x, fx, alpha = newton.armijo(fun, np.array([1.0]), np.array([-4.0]), -4.0, 0.5)
    assert alpha == 1.0 / 32
    assert x[0] == 0.875 and fx == 0.3828125
    with pytest.raises(ValueError):
        newton.armijo(fun, np.array([1.0]), np.array([1.0]), 1.0, 0.5)


def test_step_sizes_from_the_bound():
    sigma, tau = newton.step_sizes(2.0)  # 1/|A|^2 = 0.25 exactly
    assert abs(sigma * tau - 1.25 * 0.25) < 1e-15 and abs(sigma - tau) < 1e-15
    # |A| = sqrt(8), the 2-D gradient bound: 1/|A|^2 rounds to 0.12499999999999997,
    # whose decimal string truncates to 0.124
    sigma, tau = newton.step_sizes(np.sqrt(8.0))
    assert abs(sigma * tau - 1.25 * 0.124) < 1e-15
    sigma, tau = newton.step_sizes(2.0, step_factor=4.0)
    assert abs(tau / sigma - 4.0) < 1e-12


def test_newton_descends_toward_the_exact_1d_solution():
    """From a rough PDHG state every damped step lowers 0.5 |F|^2 (the Armijo
    condition with a descent direction), and 100 steps bring the iterate five
    times closer to the exact TV solution (BVLS on the dual) and the gap
    down by as much.  With the reference's CG tolerance (the squared CG
    residual below 0.5 |F|^2) the inner solves stop after a step or two near
    the solution, so the convergence is linear, not quadratic."""
    W, lam = 40, 4.0
    rng = np.random.default_rng(4)
    f = (np.repeat(rng.random(5), 8) + 0.05 * rng.standard_normal(W)).reshape(-1, 1)
    prob = newton.tv_newton_problem(f, (1, W), lam)
    base = pdhg.tv_problem(f, (1, W), lam)
    x0, y0, _, _ = pdhg.solve(base, steps="const", step_bound=0.124, term_maxiter=200, granularity=1000,
                              term_pd_gap=1e-14)
    xs, _ = tv1d.tv1d_denoise(f.ravel(), 1.0 / lam)
    xy0 = np.concatenate([x0, y0])
    sigma, tau = newton.step_sizes(2.0)
    xy, hist = newton.newton_solve(prob, xy0, tau, sigma, iters=100)
    r = [h["res_normsq"] for h in hist]
    assert all(b < a for a, b in zip(r, r[1:]))
    assert np.abs(xy[:prob.n] - xs).max() < 0.2 * np.abs(x0 - xs).max()
    assert 0 < newton.gap(prob, xy)["relgap"] < 0.2 * newton.gap(prob, xy0)["relgap"]
