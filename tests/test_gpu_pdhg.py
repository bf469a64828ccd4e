"""GPU PDHG (generic path) against the oracle PDHG.

Tolerances: FP64 iterates after a fixed iteration count agree with the FP64
oracle to 1e-9 (max abs; rounding-order differences only); FP32 iterates to
5e-4 (max abs, images in [0,1]); gaps to the same relative levels."""
import numpy as np
import pytest

from oracle import pdhg as opd

pytestmark = pytest.mark.gpu


def rects(H, W, C, seed, nrect=8):
    rng = np.random.default_rng(seed)
    img = np.full((H, W, C), 0.5)
    for _ in range(nrect):
        y0, y1 = sorted(rng.integers(0, H, 2))
        x0, x1 = sorted(rng.integers(0, W, 2))
        img[y0:y1 + 1, x0:x1 + 1, :] = rng.random(C)
    return (img + 0.1 * rng.standard_normal((H, W, C))).reshape(-1, C)


def tv_model(f, dims, lam, data="l2", tv_lbd=1.0):
    from opymize_amd.functionals import SSD, L1Norms, L1Data
    from opymize_amd.linear import GradientOp
    N, C = f.shape
    G = SSD(f, vol=lam * np.ones(C)) if data == "l2" else L1Data(f, lam)
    F = L1Norms(N, (C, len(dims)), lbd=tv_lbd)
    A = GradientOp(dims, C, scheme="forward")
    return G, F, A


@pytest.mark.parametrize("fused", [False, True])
def test_known_answer_three_iterations(gpu, fused):
    from opymize_amd.solvers import PDHG
    f = np.array([[0.0], [1.0]])
    G, F, A = tv_model(f, (1, 2), 1.0)
    for k, (xe, ye) in {1: ([0, 1 / 3], 1 / 3), 2: ([1 / 9, 4 / 9], 1 / 2), 3: ([13 / 54, 25 / 54], 5 / 9)}.items():
        s = PDHG(G, F, A)
        s.solve(steps="const", step_bound=0.25, term_maxiter=k, granularity=k, use_graph=False, fused=fused)
        x, y = s.state
        np.testing.assert_allclose(x, xe, atol=1e-15)
        np.testing.assert_allclose(y, [0, ye, 0, 0], atol=1e-15)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("steps", ["const", "precond", "adaptive"])
@pytest.mark.parametrize("data", ["l2", "l1"])
@pytest.mark.parametrize("fused", [False, True])
def test_tv_iterates_match_oracle(gpu, precision, steps, data, fused):
    from opymize_amd.solvers import PDHG
    H, W, C, lam = 20, 23, 3 if data == "l2" else 1, 4.0
    f = rects(H, W, C, 1)
    G, F, A = tv_model(f, (H, W), lam, data)
    k = 60
    s = PDHG(G, F, A)
    info = s.solve(steps=steps, precision=precision, step_bound=0.124, term_maxiter=k, granularity=k,
                   term_pd_gap=1e-30, fused=fused)
    assert (s._fused is not None) == fused
    x, y = s.state
    prob = opd.tv_problem(f, (H, W), lam, data)
    xo, yo, io, _ = opd.solve(prob, steps=steps, step_bound=0.124, term_maxiter=k, granularity=k,
                              term_pd_gap=1e-30)
    tol = 1e-9 if precision == "double" else 5e-4
    assert np.abs(x - xo).max() < tol and np.abs(y - yo).max() < tol
    rel = 1e-9 if precision == "double" else 1e-4
    for key in ("objp", "objd", "gap"):
        assert abs(info[key] - io[key]) <= rel * max(1.0, abs(io["objp"])), key
    if steps == "adaptive":
        assert abs(s.itervars["tauk"] - io["tau"]) <= 1e-9 * io["tau"]


@pytest.mark.parametrize("fused", [False, True])
def test_graph_replays_match_eager(gpu, fused):
    """The hipGraph of an iteration pair gives bitwise the eager iterates, on
    every replay, with eager work in between."""
    from opymize_amd.solvers import PDHG
    H, W = 31, 29
    f = rects(H, W, 1, 2)
    res = {}
    for use_graph in (False, True):
        G, F, A = tv_model(f, (H, W), 5.0)
        s = PDHG(G, F, A)
        outs = []

        def cb(state, info):
            outs.append((state[0].copy(), info["gap"]))

        s.solve(steps="adaptive", step_bound=0.124, term_maxiter=40, granularity=10, term_pd_gap=1e-30,
                use_graph=use_graph, cbfun=cb, fused=fused)
        res[use_graph] = outs
    assert len(res[True]) == len(res[False]) == 4
    for (xa, ga), (xb, gb) in zip(res[False], res[True]):
        np.testing.assert_array_equal(xa, xb)
        assert ga == gb


@pytest.mark.parametrize("fused", [False, True])
def test_converges_and_certifies(gpu, fused):
    from opymize_amd.solvers import PDHG
    H, W, lam = 48, 40, 8.0
    f = rects(H, W, 1, 3)
    G, F, A = tv_model(f, (H, W), lam)
    s = PDHG(G, F, A)
    info = s.solve(steps="adaptive", precision="float", step_bound=0.124, term_maxiter=20000, granularity=50,
                   term_pd_gap=1e-5, fused=fused)
    assert abs(info["relgap"]) < 1e-5 and info["iter"] < 20000
    prob = opd.tv_problem(f, (H, W), lam)
    xo, yo, io, _ = opd.solve(prob, steps="adaptive", step_bound=0.124, term_maxiter=20000, granularity=50,
                              term_pd_gap=1e-9)
    x, _ = s.state
    bound = np.sqrt(2 * info["gap"] / lam) + np.sqrt(2 * max(io["gap"], 0) / lam)
    assert np.linalg.norm(x - xo) <= bound * 1.01 + 1e-6


def test_blockop_splitsum_problem(gpu):
    """A BlockOp / SplitSum composition through the generic path."""
    from opymize_amd.functionals import SSD, L1Norms, SplitSum, IndicatorFct, ZeroFct
    from opymize_amd.linear import BlockOp, GradientOp, IdentityOp
    from opymize_amd.solvers import PDHG
    from oracle import ops, prox
    import scipy.sparse as sp
    H, W = 9, 11
    N = H * W
    f = rects(H, W, 1, 4)
    G = SplitSum([SSD(f, vol=np.array([3.0])), ZeroFct(N)])
    F = SplitSum([L1Norms(N, (1, 2), 1.0), IndicatorFct(N, c1=0.0)])
    A = BlockOp([[GradientOp((H, W), 1, scheme="forward"), 0], [IdentityOp(N), IdentityOp(N)]])
    s = PDHG(G, F, A)
    s.solve(steps="precond", term_maxiter=50, granularity=50, term_pd_gap=1e-30)
    x, y = s.state
    Am = sp.bmat([[ops.grad_matrix((H, W), 1), None], [sp.identity(N), sp.identity(N)]])

    def gprox(v, t):
        t = np.asarray(t)
        out = v.copy()
        out[:N] = prox.ssd_prox(v[:N], f, np.array([3.0]), t[:N])
        return out

    def fcprox(v, s_):
        out = v.copy()
        out[:2 * N] = prox.l1norms_proj(v[:2 * N], N, 1, 2, 1.0)
        s_ = np.asarray(s_)
        out[2 * N:] = v[2 * N:] - s_[2 * N:] * 0.0  # conj of indicator(x = 0) is 0: prox = identity
        return out

    z = lambda *_: (0.0, 0.0)
    prob = opd.Problem(Am, gprox, fcprox, z, z, z, z)
    xo, yo, _, _ = opd.solve(prob, steps="precond", term_maxiter=50, granularity=10 ** 9)
    assert np.abs(x - xo).max() < 1e-9 and np.abs(y - yo).max() < 1e-9
