"""Prox maps and projections in FP64.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py)."""
import numpy as np


def blocks_colmajor(x, N, M1, M2):
    """Per-point (M1 x M2) blocks stored column by column (entry (r, c) at
    r + c*M1), the indexing of the reference's GPU kernel proj.cu:11-21."""
    return x.reshape(N, M2, M1).transpose(0, 2, 1)


def l1norms_proj(x, N, M1, M2, lbd, matrixnorm="frobenius"):
    """Projection of every block onto {|x_i| <= lbd} (reference proj.py:87-168, proj.cu:3-107).
    Frobenius: radial scaling; spectral: singular values clipped at lbd."""
    x = np.array(x, dtype=np.float64).reshape(N, M1 * M2)
    if matrixnorm == "frobenius" or M1 == 1 or M2 == 1:
        n = np.sqrt((x ** 2).sum(axis=1))
        ext = n > lbd
        x[ext] *= (lbd / n[ext])[:, None]
        return x.ravel()
    B = blocks_colmajor(x, N, M1, M2)
    U, s, Vt = np.linalg.svd(B, full_matrices=False)
    s = np.minimum(s, lbd)
    P = np.einsum("nik,nk,nkj->nij", U, s, Vt)
    return P.transpose(0, 2, 1).reshape(N, -1).ravel()


def block_norms(x, N, M1, M2, which):
    """Frobenius / nuclear / spectral norms of the blocks (reference l1norms.py:8-32)."""
    if which == "frobenius" or M1 == 1 or M2 == 1:
        return np.sqrt((np.asarray(x).reshape(N, -1) ** 2).sum(axis=1))
    s = np.linalg.svd(blocks_colmajor(np.asarray(x, dtype=np.float64), N, M1, M2), compute_uv=False)
    return s.sum(axis=1) if which == "nuclear" else s.max(axis=1)


def reduced_monic_cubic(a, b):
    """Real root of y^3 + a y + b = 0 (reference tools/__init__.py:14-58, soln=0)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    x = np.zeros_like(a)
    z = a == 0
    x[z] = np.cbrt(-b[z])
    nz = ~z
    Q, R = a[nz] / 3.0, -b[nz] / 2.0
    Q3 = Q ** 3
    D = Q3 + R ** 2
    r = np.zeros_like(Q)
    t = D <= 0
    r[t] = 2 * np.sqrt(-Q[t]) * np.cos(np.arccos(R[t] / np.sqrt(-Q3[t])) / 3.0)
    o = ~t
    A = np.cbrt(np.abs(R[o]) + np.sqrt(D[o])) * np.sign(R[o])
    B = np.zeros_like(A)
    B[A != 0] = -Q[o][A != 0] / A[A != 0]
    r[o] = A + B
    x[nz] = r
    return x


def quadepi_proj(x, N, M, alph, shift=None, lbd=np.inf):
    """Projection onto epi{0.5*alph*|x1 - s1|^2 + s2}, optionally truncated at
    |x1 - s1| <= lbd (reference QuadEpiProj._call_cpu, proj.py:345-374)."""
    y = np.array(x, dtype=np.float64).reshape(N, M + 1)
    if shift is not None:
        y -= np.asarray(shift).reshape(N, M + 1)
    xn = np.linalg.norm(y[:, :-1], axis=-1)
    c0 = xn == 0
    if lbd < np.inf:
        with np.errstate(divide="ignore", invalid="ignore"):
            c1 = -xn / (lbd * alph) + 0.5 * alph * lbd ** 2 + 1 / alph > y[:, -1]
            c2 = 0.5 * alph * lbd ** 2 > y[:, -1]
            m = (xn > lbd) & (~c2)
            y[m, :-1] *= (lbd / xn[m])[:, None]
            m = c2 & (~c1)
            y[m, :-1] *= (lbd / xn[m])[:, None]
            y[m, -1] = 0.5 * alph * lbd ** 2
    else:
        c1 = np.ones(N, dtype=bool)
    m = c1 & (~c0) & (0.5 * alph * xn ** 2 > y[:, -1])
    a2 = 2.0 / alph ** 2
    yn = reduced_monic_cubic(a2 * (1 - y[m, -1] * alph), -a2 * xn[m])
    y[m, :-1] *= (yn / xn[m])[:, None]
    y[m, -1] = 0.5 * alph * yn ** 2
    y[c0 & (y[:, -1] < 0), -1] = 0.0
    if shift is not None:
        y += np.asarray(shift).reshape(N, M + 1)
    return y.ravel()


def quadepi_proj_bruteforce(x1, x2, alph, lbd=np.inf):
    """Independent check: by radial symmetry the projection of (x1, x2) onto
    epi{0.5 alph |y|^2, |y| <= lbd} is (r x1/|x1|, max(x2, 0.5 alph r^2)) with
    r minimising (r - |x1|)^2 + max(0, 0.5 alph r^2 - x2)^2 on [0, lbd]."""
    from scipy.optimize import minimize_scalar
    x1 = np.asarray(x1, dtype=np.float64)
    xn = np.linalg.norm(x1)
    hi = min(lbd, xn) if np.isfinite(lbd) else xn
    phi = lambda r: (r - xn) ** 2 + max(0.0, 0.5 * alph * r * r - x2) ** 2
    if hi <= 0:
        r = 0.0
    else:
        res = minimize_scalar(phi, bounds=(0.0, hi), method="bounded", options={"xatol": 1e-13, "maxiter": 2000})
        r = res.x
        if phi(hi) < phi(r):
            r = hi
    y1 = x1 * (r / xn) if xn > 0 else x1 * 0
    return y1, max(x2, 0.5 * alph * r * r)


def _tau_nc(tau, N, C):
    t = np.asarray(tau, dtype=np.float64)
    return np.full((N, C), float(t.ravel()[0])) if t.size == 1 else t.reshape(N, C)


def ssd_prox(x, f, vol, tau, mask=None):
    """(x + tau vol f)/(1 + tau vol) on unmasked points (reference quad.py:45-55,
    ShiftScaleOp affine.py:164-169 with a = 1/(1+tau vol), b = tau vol)."""
    N, C = f.shape
    tv = _tau_nc(tau, N, C) * vol[None, :]
    sh = f
    if mask is not None:
        m = np.asarray(mask, dtype=np.float64)[:, None]
        tv, sh = tv * m, f * m
    a = 1.0 / (1.0 + tv)
    return (a * (x.reshape(N, C) + tv * sh)).ravel()


def l1data_prox(x, f, lbd, tau):
    """argmin 0.5|y - x|^2 + tau*sum lbd|y - f| = f + shrink(x - f, tau*lbd)."""
    N, C = f.shape
    d = x.reshape(N, C) - f
    thr = _tau_nc(tau, N, C) * lbd[None, :]
    return (f + np.sign(d) * np.maximum(np.abs(d) - thr, 0.0)).ravel()


def l1data_conj_prox(z, f, lbd, tau):
    """prox of G*(z) = sum z f + indicator(|z| <= lbd): clip(z - tau f, -lbd, lbd)."""
    N, C = f.shape
    return np.clip(z.reshape(N, C) - _tau_nc(tau, N, C) * f, -lbd[None, :], lbd[None, :]).ravel()


def epigraph_constraints(I, J, v, b, j, i):
    """Rows (v[l], -1) and right-hand sides b[i, l] for l = J[j, k] with I[i, l],
    in the order of J[j] (reference proj.py:376-386, proj.cu:800-808)."""
    ls = [int(l) for l in J[j] if I[i, l]]
    A = -np.ones((len(ls), v.shape[1] + 1))
    A[:, :-1] = v[ls]
    return A, np.asarray(b[i, ls], dtype=np.float64)


def _proj_polyhedron_bruteforce(x, A, rhs, tol=1e-9):
    """argmin |y - x| s.t. A y <= rhs by enumerating active sets of at most
    len(x) linearly independent rows: for each set S, y = x - A_S^T mu with
    A_S y = rhs_S; a KKT point has mu >= 0 and A y <= rhs."""
    from itertools import combinations
    D = x.size
    if A.shape[0] == 0 or np.all(A @ x <= rhs + tol):
        return x.copy()
    best, bestd = None, np.inf
    for m in range(1, min(D, A.shape[0]) + 1):
        for S in combinations(range(A.shape[0]), m):
            N = A[list(S)]
            G = N @ N.T
            if np.linalg.matrix_rank(G, tol=1e-10) < m:
                continue
            mu = np.linalg.solve(G, N @ x - rhs[list(S)])
            if np.any(mu < -tol):
                continue
            y = x - N.T @ mu
            if np.all(A @ y <= rhs + tol * (1 + np.abs(rhs))):
                d = np.sum((y - x) ** 2)
                if d < bestd:
                    best, bestd = y, d
    return best


def epigraph_proj(x, I, J, v, b):
    """Projection of every x[j,i] onto the polyhedral epigraph
    {y : <v[l], y[:-1]> - y[-1] <= b[i,l], l in J[j], I[i,l]}
    (reference proj.py:388-470: the exact QP solution its CPU path obtains
    with MOSEK and its GPU path with an active-set method)."""
    I = np.asarray(I, dtype=bool)
    v = np.asarray(v, dtype=np.float64)
    nfuns = I.shape[0]
    nregions = J.shape[0]
    D = v.shape[1] + 1
    x = np.array(x, dtype=np.float64).reshape(nregions, nfuns, D)
    y = np.empty_like(x)
    for j in range(nregions):
        for i in range(nfuns):
            A, rhs = epigraph_constraints(I, J, v, b, j, i)
            y[j, i] = _proj_polyhedron_bruteforce(x[j, i], A, rhs)
    return y.ravel()


def epigraph_ind(x, I, J, v, b):
    """(0, max(0, largest constraint violation)) (reference epigraph.py:42-47)."""
    I = np.asarray(I, dtype=bool)
    nfuns, nregions, D = I.shape[0], J.shape[0], np.asarray(v).shape[1] + 1
    x = np.asarray(x, dtype=np.float64).reshape(nregions, nfuns, D)
    inf = 0.0
    for j in range(nregions):
        for i in range(nfuns):
            A, rhs = epigraph_constraints(I, J, np.asarray(v, dtype=np.float64), b, j, i)
            if A.shape[0]:
                inf = max(inf, float(np.max(A @ x[j, i] - rhs)))
    return 0.0, inf


def epigraph_pieces(I, If, J, v, b):
    """Affine pieces (g, c) with g.p + c through the ndim+1 points of every
    face If[i][j] (indices into J[j]); list over (j, i) in C order
    (reference epigraph.py:90-103)."""
    v = np.asarray(v, dtype=np.float64)
    ndim = v.shape[1]
    out = []
    for j in range(J.shape[0]):
        for i in range(np.asarray(I).shape[0]):
            eqs = []
            for face in np.asarray(If[i][j]).reshape(-1, ndim + 1):
                P = v[J[j][face]]
                f = np.asarray(b, dtype=np.float64)[i, J[j][face]]
                # [P 1] (g, c) = f
                M = np.hstack([P, np.ones((ndim + 1, 1))])
                eqs.append(np.linalg.solve(M, f))
            out.append(np.array(eqs).reshape(-1, ndim + 1))
    return out


def epigraph_supp(x, I, If, J, v, b):
    """(sum_ji xt * max_faces (g.x[:-1]/xt + c), max(0, -checks_j x)) with
    xt = max(eps, |x[-1]|) and checks_j the inverse transpose of the rows
    (v[J[j,k]], -1), k <= ndim (reference epigraph.py:104-126)."""
    v = np.asarray(v, dtype=np.float64)
    nfuns, nregions, D = np.asarray(I).shape[0], J.shape[0], v.shape[1] + 1
    x = np.asarray(x, dtype=np.float64).reshape(nregions, nfuns, D)
    pieces = epigraph_pieces(I, If, J, v, b)
    rows = -np.ones((nregions, D, D))
    rows[:, :, :-1] = v[J[:, :D]]
    checks = np.linalg.inv(rows).transpose(0, 2, 1)
    val, inf = 0.0, 0.0
    for j in range(nregions):
        for i in range(nfuns):
            xi = x[j, i]
            inf = max(inf, float(np.max(-(checks[j] @ xi))))
            xt = max(np.spacing(1), abs(xi[-1]))
            eq = pieces[j * nfuns + i]
            if eq.shape[0]:
                val += xt * float(np.max(eq[:, :-1] @ (xi[:-1] / xt) + eq[:, -1]))
    return val, max(0.0, inf)


def epigraph_supp_lp(x, I, J, v, b):
    """sum_ji max <y, x[j,i]> over the epigraph polyhedron, by linear
    programming (scipy.optimize.linprog); the definition of EpigraphSupp
    (reference epigraph.py:55-69).  Finite where x[j,i,-1] < 0 and
    x[j,i,:-1]/|x[j,i,-1]| lies in the convex hull of the region's points."""
    from scipy.optimize import linprog
    I = np.asarray(I, dtype=bool)
    v = np.asarray(v, dtype=np.float64)
    nfuns, nregions, D = I.shape[0], J.shape[0], v.shape[1] + 1
    x = np.asarray(x, dtype=np.float64).reshape(nregions, nfuns, D)
    tot = 0.0
    for j in range(nregions):
        for i in range(nfuns):
            A, rhs = epigraph_constraints(I, J, v, b, j, i)
            r = linprog(-x[j, i], A_ub=A, b_ub=rhs, bounds=[(None, None)] * D, method="highs")
            assert r.status == 0, r.message
            tot += -r.fun
    return tot
