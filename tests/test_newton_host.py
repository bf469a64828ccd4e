"""Host logic of the quasi-Newton continuation (SemismoothQuasinewton):
the line search and the Broyden iteration against scipy, whose broyden1 the
reference calls (semismooth_newton.py:461-462; scipy 1.15 here).  No GPU:
the line search is pure host code, and the Broyden iteration runs with its
three device vector kernels (dot, axpby, copy) stood in by numpy on CPU
tensors, which checks the control flow and scalar arithmetic around them."""
import numpy as np
import pytest
import scipy.optimize
from scipy.optimize._linesearch import scalar_search_armijo as scipy_armijo

from oracle import newton, pdhg


def test_scalar_search_armijo_matches_scipy():
    """The same trial steps and result as scipy's, bit for bit, over cubics
    that take each path: the full step, the quadratic step, the cubic loop,
    and no acceptable step above the minimum (None)."""
    from opymize_amd.solvers.semismooth_newton import _scalar_search_armijo
    rng = np.random.default_rng(0)
    ncalls = set()
    for _ in range(200):
        c = rng.uniform(0.0, 0.8)
        k = 10 ** rng.uniform(-1, 3)
        phi0 = rng.uniform(0.5, 2.0)
        phi = lambda s: phi0 - 2 * phi0 * s + k * (s - c) ** 2 * s
        calls_a, calls_b = [], []
        a = _scalar_search_armijo(lambda s: calls_a.append(s) or phi(s), phi0, -phi0, amin=1e-2)
        b = scipy_armijo(lambda s: calls_b.append(s) or phi(s), phi0, -phi0, amin=1e-2)
        assert a == b and calls_a == calls_b
        ncalls.add(min(len(calls_a), 3))
        ncalls.add("none" if a[0] is None else "step")
    assert ncalls == {1, 2, 3, "none", "step"}


def _patched(monkeypatch):
    import torch
    import opymize_amd.solvers.semismooth_newton as sn
    monkeypatch.setattr(sn, "_dot", lambda a, b: float(np.dot(a.numpy(), b.numpy())))
    monkeypatch.setattr(sn, "_axpby", lambda x, y, z, a, b: z.copy_(torch.from_numpy(a * x.numpy() + b * y.numpy())))
    monkeypatch.setattr(sn, "_copy", lambda s, d: d.copy_(s))
    return sn


def test_broyden_steps_match_scipy(monkeypatch):
    import torch
    sn = _patched(monkeypatch)
    W, lam = 40, 4.0
    rng = np.random.default_rng(4)
    f = (np.repeat(rng.random(5), 8) + 0.05 * rng.standard_normal(W)).reshape(-1, 1)
    prob = newton.tv_newton_problem(f, (1, W), lam)
    x0, y0, _, _ = pdhg.solve(pdhg.tv_problem(f, (1, W), lam), steps="const", step_bound=0.124, term_maxiter=200,
                              granularity=1000, term_pd_gap=1e-14)
    sigma, tau = newton.step_sizes(2.0)
    xy0 = np.concatenate([x0, y0])
    Fn = lambda v: newton.residual(prob, v, tau, sigma)[0]
    F = lambda x, out: out.copy_(torch.from_numpy(Fn(x.numpy())))
    # the same iterates up to the roundings of d = v/<df, v> (a reciprocal product here)
    # and of the dot products, which the line search's comparisons then amplify
    for it, tol in ((1, 1e-15), (2, 1e-15), (3, 1e-11), (5, 1e-8)):
        a = sn.broyden1(F, torch.from_numpy(xy0.copy()), it).numpy()
        b = scipy.optimize.broyden1(Fn, xy0.copy(), iter=it)
        assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max()), it
