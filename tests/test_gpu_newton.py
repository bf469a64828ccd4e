"""GPU semismooth Newton continuation (reference solvers/semismooth_newton.py)
against the oracle (oracle/newton.py): the new kernels (FP64 dot, axpby, the
prox Jacobians), the Newton system and its adjoint, the conjugate gradients
and whole Newton runs.

Tolerances: FP64 single-kernel results to 1e-12 relative; CG and Newton
iterates to 1e-8 (the device reductions sum in another order than numpy, and
CG amplifies the difference by its condition number); FP32 dot products to
1e-6 relative (FP64 accumulation of FP32 products)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import newton as onw
from oracle import pdhg as opd
from oracle import tv1d

pytestmark = pytest.mark.gpu


def dev(a, dtype=np.float64):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


def host(t):
    return t.cpu().numpy().astype(np.float64)


def rects(H, W, C, seed, nrect=6):
    rng = np.random.default_rng(seed)
    img = np.full((H, W, C), 0.5)
    for _ in range(nrect):
        y0, y1 = sorted(rng.integers(0, H, 2))
        x0, x1 = sorted(rng.integers(0, W, 2))
        img[y0:y1 + 1, x0:x1 + 1, :] = rng.random(C)
    return (img + 0.05 * rng.standard_normal((H, W, C))).reshape(-1, C)


def tv_model(f, dims, lam, tv_lbd=1.0):
    from opymize_amd.functionals import SSD, L1Norms
    from opymize_amd.linear import GradientOp
    N, C = f.shape
    return (SSD(f, vol=lam * np.ones(C)), L1Norms(N, (C, len(dims)), lbd=tv_lbd),
            GradientOp(dims, C, scheme="forward"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dot_and_axpby(gpu, dtype):
    import torch
    from opymize_amd.solvers.semismooth_newton import _axpby, _dot
    rng = np.random.default_rng(0)
    n = 1000003
    x = rng.standard_normal(n).astype(dtype)
    y = rng.standard_normal(n).astype(dtype)
    xd, yd = dev(x, dtype), dev(y, dtype)
    want = x.astype(np.float64).dot(y.astype(np.float64))
    assert abs(_dot(xd, yd) - want) <= (1e-12 if dtype == np.float64 else 1e-6) * np.abs(x).dot(np.abs(y))
    zd = torch.empty_like(xd)
    _axpby(xd, yd, zd, 0.75, -1.5)
    tol = 1e-15 if dtype == np.float64 else 1e-6
    np.testing.assert_allclose(host(zd), 0.75 * x.astype(np.float64) - 1.5 * y.astype(np.float64), rtol=tol,
                               atol=tol)
    _axpby(xd, yd, yd, 1.0, 2.0)  # in place on the second operand
    np.testing.assert_allclose(host(yd), x.astype(np.float64) + 2.0 * y.astype(np.float64), rtol=tol, atol=tol)
    assert _dot(xd[:0], xd[:0]) == 0.0


def test_prox_jacobians_match_oracle(gpu):
    """Each prox map called with jacobian=True returns the Jacobian at its
    input, as the oracle's matrices, also out of place and with add."""
    import torch
    from opymize_amd.operators import L1NormsProj, PosProj, NegProj, SplitOp, ConstrainOp
    from opymize_amd.functionals import SSD
    rng = np.random.default_rng(1)
    N, M, lbd = 3001, (2, 2), 0.9
    x = rng.standard_normal(N * 4) * 0.55
    v = rng.standard_normal(N * 4)
    J = onw.l12proj_jacobian(x, N, 4, lbd)
    assert 0.2 < (np.sqrt((x.reshape(N, 4) ** 2).sum(1)) > lbd).mean() < 0.8
    op = L1NormsProj(N, M, lbd, "frobenius")
    xd = dev(x)
    jac = op(xd, jacobian=True)  # in place: xd now holds the projection
    np.testing.assert_allclose(host(xd), opd.prox.l1norms_proj(x, N, 2, 2, lbd), rtol=1e-14, atol=1e-15)
    yd = torch.empty_like(xd)
    jac(dev(v), yd)
    np.testing.assert_allclose(host(yd), J @ v, rtol=1e-12, atol=1e-13)
    yd2 = dev(np.ones(N * 4))
    jac(dev(v), yd2, add=True)
    np.testing.assert_allclose(host(yd2), 1.0 + J @ v, rtol=1e-12, atol=1e-13)
    vd = dev(v)
    jac(vd)  # in place
    np.testing.assert_allclose(host(vd), J @ v, rtol=1e-12, atol=1e-13)
    # out of place: the Jacobian is taken at the input, which stays unchanged
    xd, outd = dev(x), torch.empty(N * 4, dtype=torch.float64, device="cuda")
    jac = op(xd, outd, jacobian=True)
    assert np.array_equal(host(xd), x)
    jac(dev(v), yd)
    np.testing.assert_allclose(host(yd), J @ v, rtol=1e-12, atol=1e-13)
    # sign projections
    s = rng.standard_normal(5000)
    w = rng.standard_normal(5000)
    for cls, pos in ((PosProj, True), (NegProj, False)):
        jac = cls(5000)(dev(s), jacobian=True)
        out = dev(np.zeros(5000))
        jac(dev(w), out)
        np.testing.assert_array_equal(host(out), onw.signproj_jacobian(s, pos) @ w)
    # SSD prox with volumes, a mask and an array step
    Np, C = 700, 3
    f = rng.random((Np, C))
    vol = np.array([0.5, 2.0, 4.0])
    mask = rng.random(Np) < 0.6
    for tau in (0.35, rng.random(Np * C) + 0.05):
        G = SSD(f, vol=vol, mask=mask)
        jac = G.prox(tau)(dev(rng.standard_normal(Np * C)), jacobian=True)
        u = rng.standard_normal(Np * C)
        out = dev(np.zeros(Np * C))
        jac(dev(u), out)
        np.testing.assert_allclose(host(out), onw.ssd_prox_jacobian(Np, C, vol, tau, mask) @ u, rtol=1e-14,
                                   atol=1e-15)
    # SplitOp: a block-diagonal Jacobian; ConstrainOp: the mask as a diagonal
    sp_op = SplitOp([PosProj(200), L1NormsProj(50, (1, 2), 0.5, "frobenius")])
    z = rng.standard_normal(300)
    jac = sp_op(dev(z), jacobian=True)
    Jz = sp.block_diag([onw.signproj_jacobian(z[:200], True), onw.l12proj_jacobian(z[200:], 50, 2, 0.5)])
    u = rng.standard_normal(300)
    out = dev(np.zeros(300))
    jac(dev(u), out)
    np.testing.assert_allclose(host(out), Jz @ u, rtol=1e-12, atol=1e-13)
    cmask = rng.random(100) < 0.3
    jac = ConstrainOp(cmask, rng.random((100, 2)))(dev(rng.random(200)), jacobian=True)
    out = dev(np.zeros(200))
    u = rng.standard_normal(200)
    jac(dev(u), out)
    np.testing.assert_array_equal(host(out), np.where(np.repeat(cmask, 2), 0.0, u))


def test_prox_maps_without_jacobian_refuse(gpu):
    from opymize_amd.operators import L1NormsProj, L1DataProxOp
    from opymize_amd.functionals import L1Data
    x = dev(np.random.default_rng(2).standard_normal(40))
    with pytest.raises(NotImplementedError, match="spectral"):
        L1NormsProj(10, (2, 2), 1.0, "spectral")(x, jacobian=True)
    with pytest.raises(NotImplementedError, match="Jacobian"):
        L1DataProxOp(L1Data(np.zeros((40, 1))), 0.5, False)(x, jacobian=True)


def _newton_setup(H, W, C, seed, lam=3.0, tv_lbd=0.3):
    from opymize_amd.solvers import SemismoothNewton
    f = rects(H, W, C, seed)
    G, F, A = tv_model(f, (H, W), lam, tv_lbd)
    s = SemismoothNewton(G, F, A)
    np.random.seed(seed)  # normest draws its start vector from numpy's global generator
    s.prepare_stepsizes()
    prob = onw.tv_newton_problem(f, (H, W), lam, tv_lbd)
    return s, prob, f


def test_residual_system_and_adjoint_match_oracle(gpu):
    """F(xy), M v and M^T v at a point where both prox maps have active and
    inactive parts, against the oracle's residual and Newton matrix; then
    the conjugate gradients on that system."""
    import torch
    from opymize_amd.solvers.semismooth_newton import CG_solve
    s, prob, _ = _newton_setup(11, 9, 2, 3)
    tau, sigma = s.constvars["tau"], s.constvars["sigma"]
    # tau = sigma, tau sigma = 1.25 truncate(1/|A|^2, 3) with |A| from the same power iteration
    np.random.seed(3)
    e, _ = opd.normest(prob.A)
    assert tau == pytest.approx(sigma, rel=1e-15)
    assert tau * sigma == pytest.approx(1.25 * opd.truncate(1.0 / e ** 2, 3), rel=1e-12)
    rng = np.random.default_rng(4)
    xy = np.concatenate([rng.random(prob.n), 0.25 * rng.standard_normal(prob.m)])
    xyd = dev(xy)
    res = torch.empty_like(xyd)
    r = s.res(xyd, res, subdiff=True)
    F, K, Hj = onw.residual(prob, xy, tau, sigma)
    np.testing.assert_allclose(host(res), F, rtol=1e-12, atol=1e-13)
    assert abs(r - 0.5 * F.dot(F)) <= 1e-12 * F.dot(F)
    Mo = onw.newton_matrix(prob.A, K, Hj, tau, sigma)
    v = rng.standard_normal(xy.size)
    out = torch.empty_like(xyd)
    s.M(dev(v), out)
    np.testing.assert_allclose(host(out), Mo @ v, rtol=1e-12, atol=1e-12)
    s.M.adjoint(dev(v), out)
    np.testing.assert_allclose(host(out), Mo.T @ v, rtol=1e-12, atol=1e-12)
    # CG on the regularized normal equations from a nonzero start
    x0 = 0.1 * rng.standard_normal(xy.size)
    xk = dev(x0)
    steps = CG_solve(s.M, res, xk, lbd=0.3, tol=1e-20)
    xo, so = onw.cg_solve(Mo, F, x0, lbd=0.3, tol=1e-20)
    assert abs(steps - so) <= 2
    np.testing.assert_allclose(host(xk), xo, rtol=1e-8, atol=1e-9)


def test_newton_iterates_match_oracle(gpu):
    """Eight damped Newton steps from a PDHG state: the same line-search
    steps, regularization and CG step counts as the oracle, and the same
    iterate to 1e-8."""
    from opymize_amd.solvers import SemismoothNewton
    H, W, C, lam, tv_lbd = 12, 10, 1, 4.0, 0.25
    f = rects(H, W, C, 5)
    G, F, A = tv_model(f, (H, W), lam, tv_lbd)
    base = opd.tv_problem(f, (H, W), lam, tv_lbd=tv_lbd)
    x0, y0, _, _ = opd.solve(base, steps="const", step_bound=0.124, term_maxiter=150, granularity=1000,
                             term_pd_gap=1e-14)
    s = SemismoothNewton(G, F, A)
    s.solve(continue_at=(x0, y0), term_maxiter=8, granularity=1000, term_relgap=0.0)
    assert s.iterations == 8
    prob = onw.tv_newton_problem(f, (H, W), lam, tv_lbd)
    xy, hist = onw.newton_solve(prob, np.concatenate([x0, y0]), s.constvars["tau"], s.constvars["sigma"], 8)
    gh = s.itervars["history"]
    for a, b in zip(gh, hist):
        assert a["alpha"] == b["alpha"] and a["cg_steps"] == b["cg_steps"]
        assert abs(a["lbd"] - b["lbd"]) <= 1e-12 * b["lbd"]
        assert abs(a["res_normsq"] - b["res_normsq"]) <= 1e-9 * b["res_normsq"]
    x, y = s.state
    np.testing.assert_allclose(np.concatenate([x, y]), xy, rtol=0, atol=1e-8)


def test_newton_descends_and_reports_the_gap(gpu):
    """From a rough GPU PDHG state: every step lowers 0.5 |F|^2, 100 steps
    bring x five times closer to the exact 1-D TV solution (BVLS), the
    reported gap is the oracle's at the final iterate, and a loose
    threshold ends the solve at a gap check."""
    from opymize_amd.solvers import PDHG, SemismoothNewton
    W, lam = 40, 4.0
    rng = np.random.default_rng(4)
    f = (np.repeat(rng.random(5), 8) + 0.05 * rng.standard_normal(W)).reshape(-1, 1)
    G, F, A = tv_model(f, (1, W), lam)
    p = PDHG(G, F, A)
    # the state comes back from the device at the gap checks (as in the reference): check at the end
    p.solve(steps="const", step_bound=0.124, term_maxiter=200, granularity=200, term_pd_gap=1e-14, fused=False)
    x0, y0 = p.state
    assert np.abs(y0).max() > 0.1
    xs, _ = tv1d.tv1d_denoise(f.ravel(), 1.0 / lam)
    s = SemismoothNewton(G, F, A)
    info = s.solve(continue_at=p.state, term_maxiter=100, granularity=100, term_relgap=0.0)
    r = [h["res_normsq"] for h in s.itervars["history"]]
    assert all(b < a for a, b in zip(r, r[1:]))
    x, y = s.state
    assert np.abs(x - xs).max() < 0.2 * np.abs(x0 - xs).max()
    prob = onw.tv_newton_problem(f, (1, W), lam)
    g = onw.gap(prob, np.concatenate([x, y]))
    for k in ("objp", "objd", "infeasp", "infeasd", "relgap"):
        assert abs(info[k] - g[k]) <= 1e-9 * max(1.0, abs(g[k])), k
    # thresholds: the solve ends at the first check (every 5 steps) that meets both
    s2 = SemismoothNewton(G, F, A)
    info2 = s2.solve(continue_at=p.state, term_maxiter=100, granularity=5, term_relgap=0.01, term_infeas=1e-4)
    met = [c["iter"] for c in s2.checks if abs(c["relgap"]) < 0.01 and max(c["infeasp"], c["infeasd"]) < 1e-4]
    assert [c["iter"] for c in s2.checks] == list(range(5, s2.iterations + 1, 5)), s2.checks
    assert met and met[0] == s2.iterations < 100, s2.checks
    assert info2["relgap"] == s2.checks[-1]["relgap"]


def test_newton_refuses_bad_starts_and_models(gpu):
    from opymize_amd.solvers import SemismoothNewton
    from opymize_amd.functionals import L1Data, L1Norms
    from opymize_amd.linear import GradientOp
    f = rects(6, 5, 1, 7)
    G, F, A = tv_model(f, (6, 5), 2.0)
    with pytest.raises(ValueError, match="entries"):
        SemismoothNewton(G, F, A).solve(continue_at=np.zeros(7), term_maxiter=1)
    G1 = L1Data(f, 2.0)
    with pytest.raises(NotImplementedError):
        SemismoothNewton(G1, L1Norms(30, (1, 2), 1.0), GradientOp((6, 5), 1)).solve(term_maxiter=1)


def test_quasinewton_rounds_match_scipy(gpu):
    """SemismoothQuasinewton: rounds of Broyden steps on the device against
    scipy.optimize.broyden1 on the oracle's residual (what the reference
    calls), from a rough PDHG state: the iterate after one round of 3 steps
    and two rounds of 2, and the gap reported after each round.  Broyden's
    line search branches on comparisons that amplify rounding differences,
    so the tolerance is 1e-9 relative to the largest entry."""
    from opymize_amd.solvers import SemismoothQuasinewton
    W, lam = 40, 4.0
    rng = np.random.default_rng(4)
    f = (np.repeat(rng.random(5), 8) + 0.05 * rng.standard_normal(W)).reshape(-1, 1)
    G, F, A = tv_model(f, (1, W), lam)
    x0, y0, _, _ = opd.solve(opd.tv_problem(f, (1, W), lam), steps="const", step_bound=0.124, term_maxiter=200,
                             granularity=1000, term_pd_gap=1e-14)
    prob = onw.tv_newton_problem(f, (1, W), lam)
    for gran, rounds in ((3, 1), (2, 2)):
        s = SemismoothQuasinewton(G, F, A)
        info = s.solve(continue_at=(x0, y0), granularity=gran, term_maxiter=gran * rounds, term_relgap=0.0)
        xy, gaps = onw.quasinewton_solve(prob, np.concatenate([x0, y0]), s.constvars["tau"], s.constvars["sigma"],
                                         gran, rounds)
        x, y = s.state
        scale = max(1.0, np.abs(xy).max())
        assert np.abs(np.concatenate([x, y]) - xy).max() <= 1e-9 * scale, (gran, rounds)
        assert len(s.checks) == rounds and s.iterations == gran * rounds
        for c, g in zip(s.checks, gaps):
            for k in ("objp", "objd", "infeasd"):
                assert abs(c[k] - g[k]) <= 1e-8 * max(1.0, abs(g[k])), (gran, k)
        assert info["relgap"] == s.checks[-1]["relgap"]


def test_newton_on_a_split_problem_matches_oracle(gpu):
    """A problem whose F is a SplitSum and whose A is a BlockOp: TV-L2 with
    x >= 0 written as F = (tv |.|, indicator(. >= 0)) of (grad x, x).  The
    dual prox is then a SplitOp of the ball projection and NegProj, and its
    Jacobian a block-diagonal BlockOp; five Newton steps against the oracle
    with the same blocks as scipy matrices."""
    import scipy.sparse as sp
    from opymize_amd.functionals import SSD, L1Norms, PositivityFct, SplitSum
    from opymize_amd.linear import BlockOp, GradientOp, IdentityOp
    from opymize_amd.solvers import SemismoothNewton
    from oracle import ops, prox
    H, W, lam, tvw = 9, 8, 3.0, 0.2
    N = H * W
    f = rects(H, W, 1, 11) - 0.4  # partly negative data, so that the constraint acts
    G = SSD(f, vol=np.array([lam]))
    F = SplitSum([L1Norms(N, (1, 2), lbd=tvw), PositivityFct(N)])
    A = BlockOp([[GradientOp((H, W), 1, scheme="forward")], [IdentityOp(N)]])
    rng = np.random.default_rng(12)
    xy0 = np.concatenate([rng.random(N), 0.1 * rng.standard_normal(2 * N), -0.1 * rng.random(N)])
    s = SemismoothNewton(G, F, A)
    s.solve(continue_at=xy0, term_maxiter=5, granularity=1000, term_relgap=0.0)
    tau, sigma = s.constvars["tau"], s.constvars["sigma"]
    Ao = sp.vstack([ops.grad_matrix((H, W), 1, None, None, "forward"), sp.eye(N)]).tocsr()
    lamv = np.array([lam])

    def gprox(x, t):
        return prox.ssd_prox(x, f, lamv, t), onw.ssd_prox_jacobian(N, 1, lamv, t)

    def fconjprox(y, sg):
        y1, y2 = y[:2 * N], y[2 * N:]
        val = np.concatenate([prox.l1norms_proj(y1, N, 1, 2, tvw), np.minimum(y2, 0.0)])
        jac = sp.block_diag([onw.l12proj_jacobian(y1, N, 2, tvw), onw.signproj_jacobian(y2, False)])
        return val, jac

    prob = onw.NewtonProblem(Ao, gprox, fconjprox, None, None, None, None)
    xy, hist = onw.newton_solve(prob, xy0, tau, sigma, 5)
    for a, b in zip(s.itervars["history"], hist):
        assert a["alpha"] == b["alpha"] and a["cg_steps"] == b["cg_steps"]
        assert abs(a["res_normsq"] - b["res_normsq"]) <= 1e-9 * b["res_normsq"]
    x, y = s.state
    np.testing.assert_allclose(np.concatenate([x, y]), xy, rtol=0, atol=1e-8)
    assert (xy[3 * N:] < 0).any() and (xy[N:3 * N] != 0).any()
