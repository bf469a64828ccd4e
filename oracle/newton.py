"""FP64 restatement of the semismooth Newton continuation (reference
opymize/solvers/semismooth_newton.py:14-427) on scipy sparse matrices, with
the Jacobians of the prox maps as explicit matrices.
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by tests/test_oracle_newton.py: every Jacobian against central
differences of its prox map, the Newton matrix against central differences
of the fixed-point residual, the conjugate gradients against a direct solve,
Armijo against a hand-derived backtracking, and the whole iteration against
the exact 1-D TV solution (bounded least squares)."""
import numpy as np
import scipy.sparse as sp

from oracle import pdhg, prox


# ---- Jacobians of the prox maps ----------------------------------------------------------------

def signproj_jacobian(xbar, positive):
    """PosProj / NegProj at xbar: diag(xbar > 0) / diag(xbar < 0) (reference proj.py:25, :38-40)."""
    xbar = np.asarray(xbar, dtype=np.float64).ravel()
    keep = xbar > 0 if positive else xbar < 0
    return sp.diags(keep.astype(np.float64))


def l12proj_jacobian(xbar, N, M, lbd):
    """Frobenius-ball projection at xbar (N points of M entries): identity blocks
    inside the ball, (lbd/n)(I - u u^T) with u = xbar_i/n outside (reference
    proj.py:228-275)."""
    xb = np.asarray(xbar, dtype=np.float64).reshape(N, M)
    n = np.sqrt((xb ** 2).sum(axis=1))
    blocks = []
    for i in range(N):
        if n[i] > lbd:
            u = xb[i] / n[i]
            blocks.append((lbd / n[i]) * (np.eye(M) - np.outer(u, u)))
        else:
            blocks.append(np.eye(M))
    return sp.block_diag(blocks, format="csr")


def ssd_prox_jacobian(N, C, vol, tau, mask=None):
    """SSD prox: diag(1/(1 + tau vol)) on the data points, identity elsewhere
    (reference quad.py:45-55 with ShiftScaleOp's Jacobian, affine.py:94)."""
    t = np.asarray(tau, dtype=np.float64).ravel()
    tv = np.full((N, C), t[0]) if t.size == 1 else t.reshape(N, C).copy()
    tv *= np.asarray(vol, dtype=np.float64)[None, :]
    if mask is not None:
        tv[~np.asarray(mask, dtype=bool), :] = 0.0
    return sp.diags((1.0 / (1.0 + tv)).ravel())


# ---- the iteration ------------------------------------------------------------------------------

class NewtonProblem(object):
    """min_x G(x) + F(Ax) with A sparse and the two prox maps returning
    (value, Jacobian matrix): gprox(x, tau), fconjprox(y, sigma)."""

    def __init__(self, A, gprox, fconjprox, G, F, Gc, Fc):
        self.A = sp.csr_matrix(A)
        self.AT = self.A.T.tocsr()
        self.gprox, self.fconjprox = gprox, fconjprox
        self.G, self.F, self.Gc, self.Fc = G, F, Gc, Fc
        self.n, self.m = self.A.shape[1], self.A.shape[0]


def tv_newton_problem(f, dims, lam, tv_lbd=1.0):
    """TV-L2 (SSD with volume lam) plus tv_lbd sum_i |grad x_i| (Frobenius),
    the forward-difference gradient (oracle.pdhg.tv_problem) with Jacobians."""
    base = pdhg.tv_problem(f, dims, lam, data="l2", tv_lbd=tv_lbd)
    f = np.asarray(f, dtype=np.float64)
    N, C = f.shape
    D = len(dims)
    lamv = np.broadcast_to(np.asarray(lam, dtype=np.float64), (C,)).copy()
    gprox = lambda x, t: (prox.ssd_prox(x, f, lamv, t), ssd_prox_jacobian(N, C, lamv, t))
    fconjprox = lambda y, s: (prox.l1norms_proj(y, N, C, D, tv_lbd), l12proj_jacobian(y, N, C * D, tv_lbd))
    return NewtonProblem(base.A, gprox, fconjprox, base.G, base.F, base.Gc, base.Fc)


def newton_matrix(A, K, H, tau, sigma):
    """[[I - K, tau K A^T], [-sigma H A, I - H]] (reference semismooth_newton.py:184-219)."""
    A = sp.csr_matrix(A)
    m, n = A.shape
    return sp.bmat([[sp.eye(n) - K, tau * (K @ A.T)], [-sigma * (H @ A), sp.eye(m) - H]], format="csr")


def residual(prob, xy, tau, sigma):
    """F(x, y) = (x - prox_{tau G}(x - tau A^T y), y - prox_{sigma F*}(y + sigma A x))
    and the Jacobians K, H of the two prox maps there (reference
    semismooth_newton.py:269-291)."""
    x, y = xy[:prob.n], xy[prob.n:]
    px, K = prob.gprox(x - tau * prob.AT.dot(y), tau)
    py, H = prob.fconjprox(y + sigma * prob.A.dot(x), sigma)
    return np.concatenate([x - px, y - py]), K, H


def armijo(fun, xk, p, p_gradf, fun_xk, eta=0.5, nu=0.9):
    """First alpha in 1, eta, eta^2, ... with fun(xk + alpha p) <= fun(xk) +
    alpha nu p_gradf; returns (xk + alpha p, fun there, alpha) (reference
    semismooth_newton.py:14-47)."""
    if p_gradf >= 0:
        raise ValueError("Armijo: not a descent direction")
    alpha = 1.0
    while True:
        xkp1 = xk + alpha * p
        f_xkp1 = fun(xkp1)
        if f_xkp1 <= fun_xk + alpha * nu * p_gradf:
            return xkp1, f_xkp1, alpha
        alpha *= eta
        if alpha < 1e-10:
            raise ValueError("Armijo: not a descent direction")


def cg_solve(Mmat, b, x0, lbd=0.0, maxiter=1000, tol=1e-5):
    """Conjugate gradients on (M^T M + lbd I) x = -M^T b from x0, squared
    residual tested against tol every other step (reference
    semismooth_newton.py:77-117).  Returns (x, steps)."""
    Mmat = sp.csr_matrix(Mmat)
    MT = Mmat.T.tocsr()
    op = lambda v: MT.dot(Mmat.dot(v)) + lbd * v
    x = np.array(x0, dtype=np.float64)
    r = -(op(x) + MT.dot(b))
    d = r.copy()
    rr = r.dot(r)
    k = 0
    for k in range(maxiter + 1):
        z = op(d)
        alpha = rr / d.dot(z)
        x = x + alpha * d
        r_new = r - alpha * z
        rr_new = r_new.dot(r_new)
        d = r_new + (rr_new / rr) * d
        r, rr = r_new, rr_new
        if k % 2 == 0 and rr_new < tol:
            break
    return x, k + 1


def step_sizes(op_norm, step_factor=1.0):
    """sigma, tau with tau sigma = 1.25 truncate(1/|A|^2, 3) (reference semismooth_newton.py:333-352)."""
    bnd = pdhg.truncate(1.0 / op_norm ** 2, 3) * 1.25
    sigma = np.sqrt(bnd / step_factor)
    return sigma, bnd / sigma


def newton_solve(prob, xy0, tau, sigma, iters, lbd=1.0, lbd_drop=0.1, lbd_boost=5.0):
    """`iters` damped Newton steps from xy0 (reference semismooth_newton.py:293-331,
    starting from lbd = 1, :349-352).  Returns (xy, history) with one record
    (res_normsq, alpha, rho, lbd, cg_steps) per step."""
    xy = np.array(xy0, dtype=np.float64)
    d = xy.copy()  # the first CG starts from a copy of the start (semismooth_newton.py:364), later ones
    # from the previous step's direction
    hist = []
    for _ in range(iters):
        F, K, H = residual(prob, xy, tau, sigma)
        res_normsq = 0.5 * F.dot(F)
        Mmat = newton_matrix(prob.A, K, H, tau, sigma)
        d, steps = cg_solve(Mmat, F, d, lbd=lbd, tol=res_normsq)
        Md = Mmat.dot(d)
        p_gradf = F.dot(Md)
        fun = lambda v: 0.5 * np.sum(residual(prob, v, tau, sigma)[0] ** 2)
        xy_new, new_normsq, alpha = armijo(fun, xy, d, p_gradf, res_normsq)
        rho = (new_normsq - res_normsq) / (alpha * p_gradf + alpha ** 2 * 0.5 * Md.dot(Md))
        if rho > 0.75:
            lbd *= lbd_drop
        elif rho < 0.25:
            lbd *= lbd_boost
        hist.append({"res_normsq": res_normsq, "alpha": alpha, "rho": rho, "lbd": lbd, "cg_steps": steps})
        xy = xy_new
    return xy, hist


def gap(prob, xy):
    """Objectives, infeasibilities and the relative gap as this solver writes
    it, (P - D)/max(eps, D) (reference semismooth_newton.py:395-410)."""
    x, y = xy[:prob.n], xy[prob.n:]
    gv, gi = prob.G(x)
    fv, fi = prob.F(prob.A.dot(x))
    gcv, gci = prob.Gc(-prob.AT.dot(y))
    fcv, fci = prob.Fc(y)
    objp, objd = gv + fv, -gcv - fcv
    return {"objp": objp, "objd": objd, "infeasp": gi + fi, "infeasd": gci + fci,
            "relgap": (objp - objd) / max(np.spacing(1), objd)}


def quasinewton_solve(prob, xy0, tau, sigma, granularity, rounds):
    """Rounds of scipy.optimize.broyden1(F, xy, iter=granularity) on the
    fixed-point residual, a fresh call per round, with the gap after each
    (reference semismooth_newton.py:429-500, which calls the same scipy
    routine: the oracle here is that library routine).  Returns (xy, gaps)."""
    import scipy.optimize
    xy = np.array(xy0, dtype=np.float64)
    gaps = []
    for _ in range(rounds):
        xy = scipy.optimize.broyden1(lambda v: residual(prob, v, tau, sigma)[0], xy, iter=granularity)
        gaps.append(gap(prob, xy))
    return xy, gaps
