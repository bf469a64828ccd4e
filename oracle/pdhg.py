"""FP64 restatement of PDHG(G, F, A).solve(use_gpu=False) (reference
opymize/solvers/pdhg.py:23-315) on scipy sparse matrices.
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py)."""
import numpy as np

from oracle import ops, prox


class Problem(object):
    """min_x G(x) + F(Ax) given by its sparse A, the two prox maps and the
    four (value, infeasibility) evaluations used by the gap."""

    def __init__(self, A, gprox, fconjprox, G, F, Gc, Fc, gamma=0.0, xblocks=(), yblocks=()):
        self.A = A.tocsr()
        self.AT = self.A.T.tocsr()
        self.gprox, self.fconjprox = gprox, fconjprox
        self.G, self.F, self.Gc, self.Fc = G, F, Gc, Fc
        self.gamma = gamma  # strong convexity modulus of G (accelerated steps)
        # (offset, nblocks, width): ranges of x / y whose prox is a projection onto a set that
        # couples each block of `width` entries; preconditioned steps are made constant there
        self.xblocks, self.yblocks = list(xblocks), list(yblocks)


def block_min(v, blocks):
    """Every block of `width` entries in the ranges takes its smallest positive
    entry (0 if none).  The reference's diagonal preconditioner (pdhg.py:189-200)
    gives per-element steps, but a projection onto a set that couples a block
    is the prox only for a step that is constant on the block; the smallest
    step keeps the convergence condition |Sigma^1/2 K T^1/2| <= 1."""
    v = np.array(v, dtype=np.float64)
    for off, nb, w in blocks:
        b = v[off:off + nb * w].reshape(nb, w)
        m = np.where(b > 0, b, np.inf).min(axis=1)
        m[~np.isfinite(m)] = 0.0
        v[off:off + nb * w] = np.repeat(m, w)
    return v


def tv_problem(f, dims, lam, data="l2", tv_lbd=1.0, matrixnorm="frobenius", h=None, w=None):
    """TV-L2 (SSD with volume lam) or TV-L1 (lam*|x - f|) data term plus
    tv_lbd * sum_i |grad x_i| with the forward-difference gradient.
    Blocks are (C, D) read column by column, i.e. the (N, D, C) gradient layout
    (reference examples/rof.py:69-73)."""
    f = np.asarray(f, dtype=np.float64)
    N, C = f.shape
    D = len(dims)
    A = ops.grad_matrix(dims, C, h=h, w=w, scheme="forward")
    lamv = np.broadcast_to(np.asarray(lam, dtype=np.float64), (C,)).copy()
    if data == "l2":
        gprox = lambda x, t: prox.ssd_prox(x, f, lamv, t)
        G = lambda x: (0.5 * np.einsum("ik,k->", (x.reshape(N, C) - f) ** 2, lamv), 0.0)
        Gc = lambda z: (0.5 * np.einsum("ik,k->", z.reshape(N, C) ** 2, 1.0 / lamv) + (z.reshape(N, C) * f).sum(),
                        0.0)
    else:
        gprox = lambda x, t: prox.l1data_prox(x, f, lamv, t)
        G = lambda x: (float((np.abs(x.reshape(N, C) - f) * lamv[None, :]).sum()), 0.0)
        Gc = lambda z: (float((z.reshape(N, C) * f).sum()),
                        float(max(0.0, (np.abs(z.reshape(N, C)) - lamv[None, :]).max())))
    fconjprox = lambda y, s: prox.l1norms_proj(y, N, C, D, tv_lbd, "spectral" if matrixnorm == "nuclear" else
                                               "frobenius")
    sumnorm = "nuclear" if matrixnorm == "nuclear" else "frobenius"
    ballnorm = "spectral" if matrixnorm == "nuclear" else "frobenius"
    F = lambda z: (tv_lbd * prox.block_norms(z, N, C, D, sumnorm).sum(), 0.0)
    Fc = lambda y: (0.0, float(max(0.0, (prox.block_norms(y, N, C, D, ballnorm) - tv_lbd).max())))
    return Problem(A, gprox, fconjprox, G, F, Gc, Fc, gamma=float(lamv.min()) if data == "l2" else 0.0)


def truncate(x, n):
    """Decimal-string truncation to n significant digits (reference tools/__init__.py:4-12)."""
    k = -int(np.floor(np.log10(abs(x)))) + n - 1
    s = str(abs(x))
    xs = s[:(k + 2)] if k > 0 else s[:n] + "0" * (-k)
    return np.sign(x) * float(xs)


def normest(A, tol=1.0e-4, maxits=500, rng=np.random):
    """Power iteration of |A| (reference linear/__init__.py:11-46)."""
    m, n = A.shape
    y = np.ones(m)
    y[rng.randn(m) < 0] = -1
    x = np.abs(A.T.dot(y))
    e = np.linalg.norm(x)
    if e < 1e-13:
        return e, 0
    x *= 1.0 / e
    e0, itn = 0, 0
    while abs(e - e0) > tol * e:
        e0 = e
        y = A.dot(x)
        ny = np.linalg.norm(y)
        if ny < 1e-13:
            y = rng.rand(m)
            ny = np.linalg.norm(y)
        x = A.T.dot(y)
        nx = np.linalg.norm(x)
        e = nx / ny
        x *= 1.0 / nx
        itn += 1
        if itn > maxits:
            break
    return e, itn


def pd_gap(prob, x, Ax, ATy, y):
    """(reference pdhg.py:23-44)"""
    gv, gi = prob.G(x)
    fv, fi = prob.F(Ax)
    gcv, gci = prob.Gc(-ATy)
    fcv, fci = prob.Fc(y)
    objp, objd = gv + fv, -gcv - fcv
    return {"objp": objp, "objd": objd, "infeasp": gi + fi, "infeasd": gci + fci, "gap": objp - objd,
            "relgap": (objp - objd) / max(np.spacing(1), abs(objd))}


def solve(prob, steps="const", step_bound=None, step_factor=1.0, term_pd_gap=1e-5, term_maxiter=1000,
          granularity=5000, x0=None, y0=None, term_pd_res=None, history=False, restart=None, precond_weight=1.0):
    """Iterate exactly as the reference's CPU path; returns (x, y, info, hist).
    precond_weight (preconditioned steps only, no reference counterpart)
    scales the primal steps by w and the dual steps by 1/w.
    hist (if requested) holds per-iteration (tau, sigma, alpha, resp, resd).

    restart = (beta, min_iters), accelerated steps on a strongly convex G
    only (no reference counterpart): at a test, at least min_iters
    iterations after the last restart, where |relgap| has fallen below beta
    times its value at that restart or has risen above its smallest value
    since, tau and sigma return to their initial values (x, y stay); the
    first test only records the gap.  info["restarts"] lists the iterations
    after which they happened."""
    A, AT = prob.A, prob.AT
    m, n = A.shape
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    y = np.zeros(m) if y0 is None else np.array(y0, dtype=np.float64)
    xgrad, ygrad = AT.dot(y), A.dot(x)
    info = {}
    theta = 1.0
    pd_res_mode = term_pd_res is not None
    if pd_res_mode:
        steps = "adaptive"
        eps4 = term_pd_res if isinstance(term_pd_res, tuple) else (term_pd_res,) * 4
    else:
        if not isinstance(term_pd_gap, tuple):
            term_pd_gap = (term_pd_gap, term_pd_gap)
        eps4 = (term_pd_gap[0],) * 4
    if steps == "precond":
        ysig = ops.rowwise_lp(A, 1)
        xtau = ops.rowwise_lp(AT, 1)
        eps = np.spacing(1)
        ysig[ysig > eps] = 1.0 / ysig[ysig > eps]
        xtau[xtau > eps] = 1.0 / xtau[xtau > eps]
        ysig = block_min(ysig, prob.yblocks)
        xtau = block_min(xtau, prob.xblocks)
        # primal weight w: steps w*T and Sigma/w (|Sigma^1/2 K T^1/2| is unchanged; w = 1 is the reference's)
        tau, sigma = xtau * precond_weight, ysig / precond_weight
    else:
        bnd = step_bound
        if bnd is None:
            bnd = truncate(1.0 / normest(A)[0] ** 2, 3)
        if steps == "adaptive":
            eta, Delta, alpha = 0.95, 1.5, 0.5
            tau = sigma = np.sqrt(bnd)
        else:  # constant, and the start of accelerated steps
            sigma = np.sqrt(bnd / step_factor)
            tau = bnd / sigma
    hist = []
    it = 0
    rs_ref, rs_best, rs_last, restarts = np.inf, np.inf, 0, []
    while it < term_maxiter:
        it += 1
        xn = x - tau * xgrad
        xn = prob.gprox(xn, tau)
        ygn = A.dot(xn)
        if steps == "accel":
            # Chambolle & Pock (2011), Algorithm 2 in primal-first order:
            # theta_k = 1/sqrt(1 + 2 gamma tau_k), sigma_k+1 = sigma_k/theta_k,
            # xbar = x_k+1 + theta_k (x_k+1 - x_k), tau_k+1 = theta_k tau_k
            theta = 1.0 / np.sqrt(1.0 + 2.0 * prob.gamma * tau)
            sigma = sigma / theta
        ybar = (1 + theta) * ygn - theta * ygrad
        yn = y + sigma * ybar
        yn = prob.fconjprox(yn, sigma)
        xgn = AT.dot(yn)
        if steps == "adaptive":
            resp = np.linalg.norm((x - xn) / tau - xgrad + xgn)
            resd = np.linalg.norm((y - yn) / sigma - theta * ygrad + theta * ygn)
            gp = np.linalg.norm((x - xn) / tau - xgrad)
            gd = np.linalg.norm((y - yn) / sigma - theta * ygrad + (theta + 1) * ygn)
            epsp = np.sqrt(n) * eps4[0] + eps4[1] * gp
            epsd = np.sqrt(m) * eps4[2] + eps4[3] * gd
            scale = epsp / epsd
            info.update(resp=resp, resd=resd, epsp=epsp, epsd=epsd)
            if history:
                hist.append((tau, sigma, alpha, resp, resd))
            if resp > scale * resd * Delta:
                tau, sigma, alpha = tau / (1 - alpha), sigma * (1 - alpha), alpha * eta
            if resp < scale * resd / Delta:
                tau, sigma, alpha = tau * (1 - alpha), sigma / (1 - alpha), alpha * eta
        if steps == "accel":
            tau = theta * tau
        x, xgrad, y, ygrad = xn, xgn, yn, ygn
        if it % granularity == 0:
            info["iter"] = it
            if pd_res_mode:
                done = info["resp"] < info["epsp"] and info["resd"] < info["epsd"]
            else:
                info.update(pd_gap(prob, x, ygrad, xgrad, y))
                done = abs(info["relgap"]) < term_pd_gap[0] and max(info["infeasp"], info["infeasd"]) < \
                    term_pd_gap[1]
            if done:
                break
            if steps == "accel" and restart is not None and prob.gamma > 0 and not pd_res_mode:
                g = abs(info["relgap"])
                if np.isfinite(g):
                    fire = it - rs_last >= restart[1] and ((np.isfinite(rs_ref) and g < restart[0] * rs_ref)
                                                          or g > rs_best)
                    rs_best = min(rs_best, g)
                    if not np.isfinite(rs_ref):
                        rs_ref = g
                    if fire:
                        sigma = np.sqrt(bnd / step_factor)
                        tau = bnd / sigma
                        rs_ref, rs_best, rs_last = g, np.inf, it
                        restarts.append(it)
    info["iter"] = it
    info["restarts"] = restarts
    if steps == "adaptive":
        info.update(tau=tau, sigma=sigma, alpha=alpha)
    elif steps == "accel":
        info.update(tau=tau, sigma=sigma)
    return x, y, info, hist
