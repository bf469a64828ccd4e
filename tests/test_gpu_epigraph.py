"""GPU parity of EpigraphProj and the EpigraphInd / EpigraphSupp evaluations
against the oracle (exact projection by enumeration of active sets).

Tolerances: FP64 kernels match the exact projection to 1e-10 (the active-set
method ends on the exact KKT point; only rounding differs), FP32 kernels to
2e-5 times the data scale; FP64 evaluations to 1e-12 relative."""
import numpy as np
import pytest

from oracle import prox
from tests.epi_cases import reference_proj_case, simplex_case, support_1d_case, supp_points

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 2e-5}


def dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


def host(t):
    return t.cpu().numpy().astype(np.float64)


def project(I, J, v, b, x, dtype, out=False):
    from opymize_amd.operators import EpigraphProj
    op = EpigraphProj(I, J, v, b)
    xd = dev(x.ravel(), dtype)
    if out:
        yd = dev(np.zeros(x.size), dtype)
        op(xd, yd)
        np.testing.assert_array_equal(host(xd), x.ravel().astype(dtype).astype(np.float64))
        return host(yd)
    op(xd)
    return host(xd)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_proj_reference_construction(gpu, dtype, ndim):
    for seed in range(10):
        I, J, v, b, x_on, x_in, x_out = reference_proj_case(ndim, np.random.default_rng(seed))
        sc = max(1.0, np.abs(x_on).max())
        np.testing.assert_allclose(project(I, J, v, b, x_in, dtype), x_in.ravel(), atol=TOL[dtype] * sc)
        np.testing.assert_allclose(project(I, J, v, b, x_out, dtype, out=True), x_on.ravel(),
                                   atol=TOL[dtype] * sc * 10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ndim,nregions,nfuns,ninner,drop", [
    (1, 7, 5, 6, 0.0), (1, 3, 4, 1, 0.5), (2, 5, 4, 5, 0.0), (2, 6, 3, 4, 0.4),
    (3, 4, 3, 4, 0.0), (3, 3, 2, 6, 0.3)])
def test_epigraph_proj_random(gpu, dtype, ndim, nregions, nfuns, ninner, drop):
    rng = np.random.default_rng(100 * ndim + ninner)
    I, If, J, v, b = simplex_case(ndim, nregions, nfuns, ninner, rng, drop=drop)
    x = rng.normal(0, 2, (nregions, nfuns, ndim + 1))
    x[..., -1] -= 3.0  # mostly infeasible: exercises the active-set iterations
    vt, bt = v.astype(dtype).astype(np.float64), b.astype(dtype).astype(np.float64)
    xt = x.astype(dtype).astype(np.float64)
    ref = prox.epigraph_proj(xt, I, J, vt, bt)
    got = project(I, J, v, b, x, dtype)
    sc = max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(got, ref, rtol=0, atol=TOL[dtype] * sc * 50)


def test_epigraph_proj_many_points(gpu):
    """A launch over many (region, function) pairs with several blocks."""
    rng = np.random.default_rng(7)
    I, If, J, v, b = simplex_case(2, 300, 3, 3, rng, drop=0.2)
    x = rng.normal(0, 2, (300, 3, 3))
    x[..., -1] -= 2.0
    ref = prox.epigraph_proj(x, I, J, v, b)
    got = project(I, J, v, b, x, np.float64)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9)
    # idempotent
    again = project(I, J, v, b, got.reshape(x.shape), np.float64)
    np.testing.assert_allclose(again, got, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_functionals(gpu, dtype, ndim):
    from opymize_amd.functionals import EpigraphInd, EpigraphSupp
    rng = np.random.default_rng(20 + ndim)
    I, If, J, v, b = simplex_case(ndim, 4, 3, 4, rng)
    G = EpigraphInd(I, If, J, v, b)
    S = G.conj
    assert isinstance(S, EpigraphSupp) and S.conj is G
    # indicator: infeasible points
    x = rng.normal(0, 1, (4, 3, ndim + 1))
    xt = x.astype(dtype).astype(np.float64)
    val, inf = G(dev(x.ravel(), dtype))
    _, rinf = prox.epigraph_ind(xt, I, J, v, b)
    assert val == 0.0
    np.testing.assert_allclose(inf, rinf, rtol=1e-12, atol=1e-12)
    # projection through the prox of the indicator removes the infeasibility
    xd = dev(x.ravel(), dtype)
    G.prox(0.5)(xd)
    assert G(xd)[1] <= (1e-12 if dtype == np.float64 else 1e-5)
    # support function at points where it is finite, and the LP value
    z = supp_points(J, v, 3, rng)
    zt = z.astype(dtype).astype(np.float64)
    sval, sinf = S(dev(z.ravel(), dtype))
    rval, rinf = prox.epigraph_supp(zt, I, If, J, v, b)
    # the affine pieces come from small solves on thin Delaunay faces, whose
    # conditioning (not the kernel) limits the agreement to about 1e-9
    np.testing.assert_allclose(sval, rval, rtol=1e-8, atol=1e-8)
    assert sinf == pytest.approx(rinf, abs=1e-12)
    if dtype == np.float64:
        np.testing.assert_allclose(sval, prox.epigraph_supp_lp(z, I, J, v, b), rtol=1e-8, atol=1e-8)
    # a point outside the cone of region 0 is reported infeasible
    z[0, 0, -1] = 1.0
    _, sinf = S(dev(z.ravel(), np.float64))
    assert sinf > 0
    assert sinf == pytest.approx(prox.epigraph_supp(z, I, If, J, v, b)[1], rel=1e-12)


def test_epigraph_support_1d_reference_construction(gpu):
    from opymize_amd.functionals import EpigraphSupp
    errs = []
    for n in [5, 50, 500]:
        I, If, J, v, b, x, exact = support_1d_case(n, np.random.default_rng(1))
        S = EpigraphSupp(I, If, J, v, b)
        val, inf = S(dev(x.ravel(), np.float64))
        assert inf <= 1e-12
        np.testing.assert_allclose(val, prox.epigraph_supp(x, I, If, J, v, b)[0], rtol=1e-12)
        errs.append(abs(val - exact))
    assert errs[2] < errs[1] < errs[0]


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_epigraph_pdhg_trajectory(gpu, ndim, use_graph):
    """PDHG on min_x G(x) + EpigraphSupp(x) with G the indicator of {x = z}
    and tau = sigma = 1: x_k = z, y_{k+1} = P(y_k + (2 x_{k+1} - x_k)), the
    projection being EpigraphProj as the prox of F* = EpigraphInd.  The GPU
    iterates follow the oracle's iteration (exact projections), the primal
    value is the support function at z, the dual iterates stay feasible and
    weak duality holds."""
    from opymize_amd.functionals import EpigraphSupp, IndicatorFct
    from opymize_amd.linear import IdentityOp
    from opymize_amd.solvers import PDHG
    rng = np.random.default_rng(5)
    nregions, nfuns, K = 6, 2, 40
    I, If, J, v, b = simplex_case(ndim, nregions, nfuns, 1, rng, alpha=20.0)
    n = nregions * nfuns * (ndim + 1)
    z = supp_points(J, v, nfuns, rng)
    pd = PDHG(IndicatorFct(n, c1=z.ravel()), EpigraphSupp(I, If, J, v, b), IdentityOp(n))
    pd.solve(steps="const", step_bound=1.0, term_pd_gap=1e-12, term_maxiter=K, granularity=K, use_graph=use_graph)
    y = np.zeros(n)
    xk = np.zeros(n)
    for _ in range(K):
        y = prox.epigraph_proj(y + (2 * z.ravel() - xk), I, J, v, b)
        xk = z.ravel()
    np.testing.assert_allclose(pd.itervars["xk"], z.ravel(), atol=1e-14)
    np.testing.assert_allclose(pd.itervars["yk"], y, rtol=0, atol=1e-9 * max(1.0, np.abs(y).max()))
    ref = prox.epigraph_supp(z, I, If, J, v, b)[0]
    np.testing.assert_allclose(pd.info["objp"], ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pd.info["objd"], y @ z.ravel(), rtol=1e-10, atol=1e-10)
    assert pd.info["objd"] <= pd.info["objp"] + 1e-9
    assert pd.info["infeasd"] <= 1e-9
