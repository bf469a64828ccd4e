"""Pins of the oracle's prox maps and projections (CPU only)."""
import numpy as np
import pytest

from oracle import prox


def test_l1norms_frobenius_closed_form():
    rng = np.random.default_rng(0)
    N, M1, M2, lbd = 50, 3, 2, 0.7
    x = rng.standard_normal(N * M1 * M2) * 0.6
    p = prox.l1norms_proj(x, N, M1, M2, lbd, "frobenius").reshape(N, -1)
    xr = x.reshape(N, -1)
    n = np.linalg.norm(xr, axis=1)
    for i in range(N):
        if n[i] <= lbd:
            np.testing.assert_array_equal(p[i], xr[i])
        else:
            np.testing.assert_allclose(p[i], xr[i] * lbd / n[i], rtol=1e-14)
    # idempotent
    np.testing.assert_allclose(prox.l1norms_proj(p.ravel(), N, M1, M2, lbd, "frobenius"), p.ravel(), rtol=1e-14)


@pytest.mark.parametrize("M1,M2", [(3, 2), (2, 3), (2, 2)])
def test_l1norms_spectral_projection_properties(M1, M2):
    rng = np.random.default_rng(1)
    N, lbd = 40, 0.8
    x = rng.standard_normal(N * M1 * M2)
    p = prox.l1norms_proj(x, N, M1, M2, lbd, "spectral")
    B = prox.blocks_colmajor(p, N, M1, M2)
    s = np.linalg.svd(B, compute_uv=False)
    assert s.max() <= lbd + 1e-12
    # projection: <x - p, q - p> <= 0 for feasible q (random feasible points)
    for _ in range(20):
        q = rng.standard_normal(N * M1 * M2)
        q = prox.l1norms_proj(q, N, M1, M2, lbd, "spectral")
        ip = ((x - p).reshape(N, -1) * (q - p).reshape(N, -1)).sum(axis=1)
        assert ip.max() <= 1e-10
    # blocks read column by column: an (M1 x M2) = (C, D) block of the (D, C) gradient layout
    g = np.arange(M1 * M2, dtype=np.float64)
    B0 = prox.blocks_colmajor(g, 1, M1, M2)[0]
    assert B0[1, 0] == 1.0 and B0[0, 1] == M1


def test_cubic_against_numpy_roots():
    rng = np.random.default_rng(2)
    a = rng.uniform(-5, 5, 300)
    b = rng.uniform(-5, 5, 300)
    a[:5] = 0.0
    r = prox.reduced_monic_cubic(a, b)
    np.testing.assert_allclose(r ** 3 + a * r + b, 0.0, atol=1e-9)
    for ai, bi, ri in zip(a, b, r):
        roots = np.roots([1.0, 0.0, ai, bi])
        real = roots[np.abs(roots.imag) < 1e-7].real
        assert np.min(np.abs(real - ri)) < 1e-7


def _sample_ball(rng, N, M):
    v = rng.standard_normal((N, M))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v * rng.random((N, 1)) ** (1 / M)


@pytest.mark.parametrize("seed", range(5))
def test_quadepi_reference_invariants(seed):
    # reference tests/quadepi.py:21-119, tolerance 1e-10
    rng = np.random.default_rng(seed)
    N, M = int(rng.integers(1, 10)), int(rng.integers(1, 5))
    a = 20 * rng.random()
    b = -5.0 + 10 * rng.random((N, M))
    c = -5.0 + 10 * rng.random(N)
    f = lambda x1: (0.5 * a * x1 ** 2 + b * x1).sum(axis=1) + c
    shift = np.zeros((N, M + 1))
    shift[:, :-1] = -b / a
    shift[:, -1] = -(0.5 / a * (b ** 2).sum(axis=-1) - c)
    # inside: fixed
    x = np.zeros((N, M + 1))
    x[:, :-1] = 10 * _sample_ball(rng, N, M)
    x[:, -1] = f(x[:, :-1]) + rng.random(N)
    y = prox.quadepi_proj(x.ravel(), N, M, a, shift).reshape(N, M + 1)
    assert np.linalg.norm(x - y) < 1e-10
    # outside: lands on the paraboloid
    x[:, -1] = f(x[:, :-1]) - 3 * rng.random(N)
    y = prox.quadepi_proj(x.ravel(), N, M, a, shift).reshape(N, M + 1)
    assert np.all(np.abs(f(y[:, :-1]) - y[:, -1]) < 1e-10)


@pytest.mark.parametrize("seed", range(5))
def test_quadepi_truncated_invariants(seed):
    rng = np.random.default_rng(10 + seed)
    N, M = int(rng.integers(1, 10)), int(rng.integers(1, 5))
    a = 20 * rng.random()
    lbd = 5 * rng.random()
    f = lambda x1: 0.5 * a * (x1 ** 2).sum(axis=1)
    x = np.zeros((N, M + 1))
    v = rng.standard_normal((N, M))
    v /= np.linalg.norm(v, axis=1)[:, None]
    x[:, :-1] = lbd * v
    fx = f(x[:, :-1])
    x[:, :-1] *= (1 + rng.random((N, 1)))
    x[:, -1] = fx + 3 * rng.random(N)
    y = prox.quadepi_proj(x.ravel(), N, M, a, lbd=lbd).reshape(N, M + 1)
    assert np.all(fx - y[:, -1] < 1e-10)
    assert np.abs(np.linalg.norm(y[:, :-1], axis=-1) - lbd).max() < 1e-10


@pytest.mark.parametrize("lbd", [np.inf, 0.6, 2.0])
def test_quadepi_against_bruteforce(lbd):
    rng = np.random.default_rng(20)
    N, M, a = 60, 2, 1.7
    x = rng.standard_normal((N, M + 1)) * 2
    y = prox.quadepi_proj(x.ravel(), N, M, a, lbd=lbd).reshape(N, M + 1)
    for i in range(N):
        y1, y2 = prox.quadepi_proj_bruteforce(x[i, :-1], x[i, -1], a, lbd)
        np.testing.assert_allclose(y[i, :-1], y1, atol=2e-6)
        np.testing.assert_allclose(y[i, -1], y2, atol=2e-6)


def test_ssd_prox_moreau_identity():
    # x = prox_{tG}(x) + t * prox_{G*/t}(x/t) (reference opymize/__init__.py:170-171)
    rng = np.random.default_rng(30)
    N, C = 20, 3
    f = rng.standard_normal((N, C))
    vol = rng.uniform(0.5, 3, C)
    x = rng.standard_normal(N * C)
    t = 0.37
    p = prox.ssd_prox(x, f, vol, t)
    # conj of SSD(f, vol) = SSD(-vol f, vol=1/vol) + const
    q = prox.ssd_prox(x / t, -f * vol[None, :], 1.0 / vol, 1.0 / t)
    np.testing.assert_allclose(p + t * q, x, atol=1e-13)
    # first-order optimality: (p - x) + t vol (p - f) = 0
    np.testing.assert_allclose((p - x).reshape(N, C) + t * vol[None, :] * (p.reshape(N, C) - f), 0, atol=1e-13)


def test_l1data_prox_and_moreau():
    rng = np.random.default_rng(31)
    N, C = 30, 2
    f = rng.standard_normal((N, C))
    lbd = rng.uniform(0.5, 2, C)
    x = rng.standard_normal(N * C) * 2
    t = 0.6
    p = prox.l1data_prox(x, f, lbd, t)
    # brute force per entry on a fine grid around the candidate
    for i in range(0, N * C, 7):
        k = i % C
        fi = f.ravel()[i]
        grid = np.linspace(p[i] - 0.05, p[i] + 0.05, 20001)
        obj = 0.5 * (grid - x[i]) ** 2 + t * lbd[k] * np.abs(grid - fi)
        assert abs(grid[np.argmin(obj)] - p[i]) < 1e-5
    q = prox.l1data_conj_prox(x / t, f, lbd, 1.0 / t)
    np.testing.assert_allclose(p + t * q, x, atol=1e-12)


# ---- polyhedral epigraphs -------------------------------------------------

@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_proj_reference_construction(ndim):
    """The reference's tests/epigraph.py:11-59: points above the epigraph are
    fixed, points pushed out along a facet normal return to the facet."""
    from tests.epi_cases import reference_proj_case
    for seed in range(10):
        I, J, v, b, x_on, x_in, x_out = reference_proj_case(ndim, np.random.default_rng(seed))
        np.testing.assert_allclose(prox.epigraph_proj(x_in, I, J, v, b), x_in.ravel(), atol=1e-12)
        np.testing.assert_allclose(prox.epigraph_proj(x_out, I, J, v, b), x_on.ravel(), atol=1e-9)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_proj_variational_inequality(ndim):
    """y = P(x) is feasible and <x - y, z - y> <= 0 for feasible z (the
    characterisation of the projection onto a closed convex set); checked
    against an independent solver (scipy SLSQP) as well."""
    from scipy.optimize import minimize
    from tests.epi_cases import simplex_case
    rng = np.random.default_rng(ndim)
    I, If, J, v, b = simplex_case(ndim, 2, 3, 3, rng, drop=0.3)
    x = rng.normal(0, 2, (2, 3, ndim + 1))
    y = prox.epigraph_proj(x, I, J, v, b).reshape(x.shape)
    for j in range(2):
        for i in range(3):
            A, rhs = prox.epigraph_constraints(I, J, v, b, j, i)
            assert np.all(A @ y[j, i] <= rhs + 1e-9)
            for _ in range(20):  # random feasible points: lift above the constraints
                z = rng.normal(0, 3, ndim + 1)
                z[-1] += max(0.0, np.max(A @ z - rhs)) + rng.random()
                assert (x[j, i] - y[j, i]) @ (z - y[j, i]) <= 1e-9
            r = minimize(lambda q: 0.5 * np.sum((q - x[j, i]) ** 2), x[j, i] + 0, jac=lambda q: q - x[j, i],
                         constraints=[{"type": "ineq", "fun": lambda q: rhs - A @ q, "jac": lambda q: -A}],
                         method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
            np.testing.assert_allclose(y[j, i], r.x, atol=1e-6)


def test_epigraph_ind_violation():
    from tests.epi_cases import simplex_case
    rng = np.random.default_rng(3)
    I, If, J, v, b = simplex_case(2, 3, 2, 2, rng)
    x = rng.normal(0, 1, (3, 2, 3))
    x[..., -1] = -20.0  # below every epigraph: projections land on the boundary
    y = prox.epigraph_proj(x, I, J, v, b)
    assert prox.epigraph_ind(y, I, J, v, b)[1] <= 1e-9
    y = y.reshape(3, 2, 3)
    y[1, 0, -1] -= 0.25  # lowering the height by t violates the active constraints by t
    _, inf = prox.epigraph_ind(y, I, J, v, b)
    A, rhs = prox.epigraph_constraints(I, J, v, b, 1, 0)
    assert abs(inf - max(0.0, np.max(A @ y[1, 0] - rhs))) < 1e-12
    assert inf >= 0.25 - 1e-9


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_supp_is_lp_value(ndim):
    """The faces formula equals max <y, x> over the epigraph polyhedron."""
    from tests.epi_cases import simplex_case, supp_points
    rng = np.random.default_rng(10 + ndim)
    I, If, J, v, b = simplex_case(ndim, 3, 2, 4, rng)
    x = supp_points(J, v, 2, rng)
    val, inf = prox.epigraph_supp(x, I, If, J, v, b)
    assert inf <= 1e-12
    np.testing.assert_allclose(val, prox.epigraph_supp_lp(x, I, J, v, b), rtol=1e-9, atol=1e-9)


def test_epigraph_supp_converges_to_perspective():
    """The reference's tests/epigraph.py:61-105: sampled quadratics, the error
    against the perspective of the quadratic falls as the sampling refines."""
    from tests.epi_cases import support_1d_case
    errs = []
    for n in [5, 50, 500]:
        I, If, J, v, b, x, exact = support_1d_case(n, np.random.default_rng(1))
        val, inf = prox.epigraph_supp(x, I, If, J, v, b)
        assert inf <= 1e-12
        errs.append(abs(val - exact))
    assert errs[1] < errs[0] and errs[2] < errs[1] and errs[2] < 1e-3
