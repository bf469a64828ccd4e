"""Polyhedral-epigraph test instances (shared by the CPU and GPU tests).

``reference_proj_case`` and ``support_1d_case`` rebuild the constructions of
the reference's tests/epigraph.py:11-59 and :61-105 with a seeded generator;
``simplex_case`` builds simplicial regions with interior sub-points and a
convex piecewise linear function per (region, function), faces from the
Delaunay triangulation of the region's points (the lower hull of points
lifted onto a convex quadratic)."""
import numpy as np


def reference_proj_case(ndim, rng):
    """One region, two functions over three points: function 0 uses point 0,
    function 1 uses points 1 and 2.  Returns (I, J, v, b, x_on, x_in, x_out):
    x_on lies on the boundary of each epigraph, x_in above it (a fixed point
    of the projection), x_out is x_on pushed out along the normal of the
    active facet (projects back onto x_on)."""
    npoints, nfuns = 3, 2
    I = np.array([[True, False, False], [False, True, True]])
    J = np.arange(npoints)[None, :]
    v = -2.0 + 4 * rng.random((npoints, ndim))
    b = -2.0 + 4 * rng.random((nfuns, npoints))
    x_on = np.zeros((1, nfuns, ndim + 1))
    x_on[0, 0] = -5.0 + 10 * rng.random(ndim + 1)
    x_on[0, 0, -1] = -b[0, 0] + v[0] @ x_on[0, 0, :-1]
    x_on[0, 1] = -5.0 + 10 * rng.random(ndim + 1)
    vals = [-bk + vk @ x_on[0, 1, :-1] for bk, vk in zip(b[1, 1:], v[1:])]
    k = int(np.argmax(vals))
    x_on[0, 1, -1] = vals[k]
    x_in = x_on.copy()
    x_in[0, :, -1] += 4 * rng.random(nfuns)
    x_out = x_on.copy()
    s = rng.random(2)
    x_out[0, 0, :-1] += s[0] * v[0]
    x_out[0, 0, -1] -= s[0]
    x_out[0, 1, :-1] += s[1] * v[1 + k]
    x_out[0, 1, -1] -= s[1]
    return I, J, v, b, x_on, x_in, x_out


def support_1d_case(nsub, rng, nfuns=5, nregions=2):
    """1-D quadratics a(x - p)^2 + c sampled on nsub points of two intervals
    (the interval ends first, as simplex vertices), faces between neighbours;
    x[j,i] = (-xt * p, xt) with xt < 0 and p inside interval j.  Returns
    (I, If, J, v, b, x, exact) with exact = sum_ji -xt f_i(-x/xt)."""
    fct = lambda x, a, p, c: a * (x - p) ** 2 + c
    x0 = 5 * (-1 + 2 * rng.random(nregions))
    x1 = x0 + 0.1 + 2 * rng.random(nregions)
    a = 0.1 + 4 * rng.random(nfuns)
    p = x0[0] + (-1.5 + 4 * rng.random(nfuns)) * (x1[0] - x0[0])
    c = 5 * (-1 + 2 * rng.random(nfuns))
    npoints = nregions * nsub
    I = np.ones((nfuns, npoints), dtype=bool)
    J = np.arange(npoints).reshape(nregions, nsub)
    v = np.zeros((nregions, nsub, 1))
    for j in range(nregions):
        v[j, :, 0] = np.linspace(x0[j], x1[j], nsub)
        last = J[j, -1]
        J[j, 2:] = J[j, 1:-1].copy()
        J[j, 1] = last
    v = v.reshape(npoints, 1)
    b = np.stack([fct(v[:, 0], a[i], p[i], c[i]) for i in range(nfuns)])
    faces = np.array([(0, 2), (nsub - 1, 1)] + [(k, k + 1) for k in range(2, nsub - 1)])
    If = [[faces.copy() for _ in range(nregions)] for _ in range(nfuns)]
    x = np.zeros((nregions, nfuns, 2))
    for j in range(nregions):
        x[j, :, -1] = -rng.random()
        x[j, :, 0] = -x[j, :, -1] * (x0[j] + rng.random(nfuns) * (x1[j] - x0[j]))
    exact = 0.0
    for i in range(nfuns):
        for j in range(nregions):
            xx, xt = x[j, i, 0], x[j, i, 1]
            exact += -xt * fct(-xx / xt, a[i], p[i], c[i])
    return I, If, J, v, b, x, exact


def simplex_case(ndim, nregions, nfuns, ninner, rng, drop=0.0, alpha=1.0):
    """Random simplices (vertices first in J[j]) with ninner interior points
    each (barycentric coordinates ~ Dirichlet(alpha); large alpha keeps them
    near the centroid), convex quadratic values b[i] = q_i |v - m_i|^2 + c_i,
    faces from the Delaunay triangulation of each region's points.  With
    drop > 0 a fraction of the interior points is deselected in I (vertices
    stay selected, so the faces remain valid only for drop = 0)."""
    from scipy.spatial import Delaunay
    nsub = ndim + 1 + ninner
    npoints = nregions * nsub
    v = np.zeros((npoints, ndim))
    J = np.arange(npoints).reshape(nregions, nsub)
    for j in range(nregions):
        while True:
            vert = rng.uniform(-2, 2, (ndim + 1, ndim))
            if abs(np.linalg.det(vert[1:] - vert[:1])) > 0.5:
                break
        lam = rng.dirichlet(alpha * np.ones(ndim + 1), ninner)
        v[J[j]] = np.vstack([vert, lam @ vert])
    q = rng.uniform(0.3, 2.0, nfuns)
    m = rng.uniform(-1, 1, (nfuns, ndim))
    c = rng.uniform(-1, 1, nfuns)
    b = np.stack([q[i] * ((v - m[i]) ** 2).sum(axis=1) + c[i] for i in range(nfuns)])
    I = np.ones((nfuns, npoints), dtype=bool)
    if drop > 0:
        inner = np.zeros(npoints, dtype=bool)
        inner[J[:, ndim + 1:].ravel()] = True
        I &= ~(inner[None, :] & (rng.random((nfuns, npoints)) < drop))
    If = [[None] * nregions for _ in range(nfuns)]
    for j in range(nregions):
        if ndim == 1:
            order = np.argsort(v[J[j], 0])
            faces = np.stack([order[:-1], order[1:]], axis=1)
        else:
            faces = Delaunay(v[J[j]]).simplices
        for i in range(nfuns):
            If[i][j] = faces.copy()
    return I, If, J, v, b


def supp_points(J, v, nfuns, rng):
    """x[j,i] = (-xt * p, xt), xt < 0, p a random convex combination of the
    simplex vertices of region j (where EpigraphSupp is finite)."""
    nregions = J.shape[0]
    ndim = v.shape[1]
    x = np.zeros((nregions, nfuns, ndim + 1))
    for j in range(nregions):
        vert = v[J[j, :ndim + 1]]
        lam = rng.dirichlet(np.ones(ndim + 1), nfuns)
        xt = -rng.uniform(0.2, 2.0, nfuns)
        x[j, :, :-1] = -xt[:, None] * (lam @ vert)
        x[j, :, -1] = xt
    return x
