"""Semismooth Newton continuation of PDHG on the GPU (reference:
opymize/solvers/semismooth_newton.py:14-427).

Newton's method on the fixed-point residual of PDHG,

    F(x, y) = (x - prox_{tau G}(x - tau A^T y),  y - prox_{sigma F*}(y + sigma A x)),

whose generalized Jacobian is

    M = [[I - K, tau K A^T], [-sigma H A, I - H]]

with K, H the Jacobians of the two prox maps at the current point.  Each step
solves the regularized normal equations (M^T M + lbd I) d = -M^T F by
conjugate gradients, backtracks along d until 0.5 |F|^2 decreases enough
(Armijo), and raises or lowers lbd by the ratio of actual to predicted
decrease.  It is meant to continue from a PDHG state and polish it to a
small gap.

Every vector stays on the device: the linear operator, the prox maps and
their Jacobians are HIP kernels, the vector updates are ``opy_axpby``, the
inner products FP64 reductions (``opy_dot``).  The host sees only the
scalars that steer the iteration (the CG coefficients, the Armijo test, lbd)
and the gap every ``granularity`` steps.

The reference runs this solver on the CPU only (its Jacobians have no GPU
path); here it runs on the GPU with the same steps, so its iterates follow
the reference's rule step for step.  Jacobians exist for the prox maps whose reference has one: PosProj,
NegProj, ConstOp, ConstrainOp, ShiftScaleOp, L1NormsProj (Frobenius), SSD's
prox, SplitOp of those, and every LinOp (itself).
"""
import logging
import math

import numpy as np
import torch

from opymize_amd import LinOp, Variable, _hip
from opymize_amd.linear import normest
from opymize_amd.tools import truncate
from opymize_amd.tools.util import GracefulInterruptHandler


_dot_out = {}  # one FP64 result slot per device


def _dot(a, b):
    """<a, b> in FP64 on the device, returned to the host."""
    if b.numel() != a.numel():
        raise ValueError("dot: %d and %d entries" % (a.numel(), b.numel()))
    dev = a.device.index if a.device.index is not None else torch.cuda.current_device()
    out = _dot_out.get(dev)
    if out is None:
        out = torch.zeros(1, dtype=torch.float64, device=a.device)
        _dot_out[dev] = out
    _hip.call("dot", a.dtype, _hip.ptr(a), _hip.ptr(b), a.numel(), _hip.ptr(_hip.Workspace.get()),
              _hip.ptr(out), _hip.stream())
    return float(out.item())


def _axpby(x, y, z, a, b):
    """z = a*x + b*y on the device (z may be x or y)."""
    if y.numel() != x.numel() or z.numel() != x.numel():
        raise ValueError("axpby: %d, %d and %d entries" % (x.numel(), y.numel(), z.numel()))
    _hip.call("axpby", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(z), x.numel(), float(a), float(b),
              _hip.stream())


def _copy(src, dst):
    _hip.call("copy", src.dtype, _hip.ptr(src), _hip.ptr(dst), src.numel(), _hip.stream())


def armijo(fun, xk, xkp1, p, p_gradf, fun_xk, eta=0.5, nu=0.9):
    """Backtracking line search: the first alpha in 1, eta, eta^2, ... with
    fun(xk + alpha p) <= fun(xk) + alpha nu p_gradf; sets xkp1 = xk + alpha p
    and returns (fun(xkp1), alpha) (reference: semismooth_newton.py:14-47)."""
    if p_gradf >= 0:
        raise Exception("Armijo: the direction does not descend (slope %g)" % p_gradf)
    alpha = 1.0
    while True:
        _axpby(xk, p, xkp1, 1.0, alpha)
        f_xkp1 = fun(xkp1)
        if f_xkp1 <= fun_xk + alpha * nu * p_gradf:
            break
        alpha *= eta
        if alpha < 1e-10:
            raise Exception("Armijo: no sufficient decrease down to step %g" % alpha)
    return f_xkp1, alpha


class SymRegOp(LinOp):
    """M^T M + lbd I for the given M (reference: semismooth_newton.py:49-75)."""

    def __init__(self, M, lbd=1.0, tmp=None):
        LinOp.__init__(self)
        self.M = M
        self.lbd = lbd
        self.x = M.x
        self.y = M.x
        self.tmp = tmp
        self.adjoint = self

    def _call_gpu(self, x, y=None, add=False):
        assert y is not None
        if self.tmp is None or self.tmp.numel() != self.M.y.size or self.tmp.dtype != x.dtype:
            self.tmp = torch.empty(self.M.y.size, dtype=x.dtype, device=x.device)
        self.M(x, self.tmp)
        self.M.adjoint(self.tmp, y, add=add)
        if self.lbd != 0:
            _axpby(y, x, y, 1.0, self.lbd)


def CG_solve(A, b, xk, lbd=0, tmpvars=None, maxiter=int(1e3), tol=1e-5):
    """Solve A x + b = 0 in the least-squares sense by conjugate gradients on
    (A^T A + lbd I) x = -A^T b, starting from and writing to xk; the squared
    residual is tested against tol every other step (reference:
    semismooth_newton.py:77-117).  Returns the number of steps taken."""
    if tmpvars is not None:
        rk, rkp1, z, dk = tmpvars
    else:
        rk, rkp1, z, dk = [torch.empty_like(xk) for _ in range(4)]
    ATb = z
    modA = SymRegOp(A, lbd=lbd)
    A.adjoint(b, ATb)
    modA(xk, rk)
    _axpby(rk, ATb, rk, -1.0, -1.0)  # rk = -(A^T A + lbd) xk - A^T b
    _copy(rk, dk)
    rr = _dot(rk, rk)
    k = 0
    for k in range(maxiter + 1):
        modA(dk, z)
        alpha = rr / _dot(dk, z)
        _axpby(xk, dk, xk, 1.0, alpha)
        _axpby(rk, z, rkp1, 1.0, -alpha)
        rr_new = _dot(rkp1, rkp1)
        beta = rr_new / rr
        _axpby(rkp1, dk, dk, 1.0, beta)
        _copy(rkp1, rk)
        rr = rr_new
        if k % 2 == 0:
            logging.debug("CG #{:03d}: {: 9.6g}".format(k, rr_new))
            if rr_new < tol:
                break
    return k + 1


class SemismoothNewtonSystem(LinOp):
    """[[I - K, tau K A^T], [-sigma H A, I - H]] for the prox Jacobians K, H
    set by the residual evaluation (reference: semismooth_newton.py:184-219)."""

    def __init__(self, A, tau, sigma):
        LinOp.__init__(self)
        self.A = A
        self.tau = tau
        self.sigma = sigma
        self.x = Variable((A.x.size,), (A.y.size,))
        self.y = self.x
        self.xtmp = None
        self.K = None
        self.H = None
        self.adjoint = SemismoothNewtonSystemAdjoint(self)

    def _tmp(self, like):
        if self.xtmp is None or self.xtmp.dtype != like.dtype or self.xtmp.device != like.device:
            self.xtmp = torch.empty(self.x.size, dtype=like.dtype, device=like.device)
        return [v.reshape(-1) for v in self.x.vars(self.xtmp)]

    def _call_gpu(self, x, y=None, add=False):
        assert y is not None
        assert add is False
        x1tmp, x2tmp = self._tmp(x)
        x1, x2 = [v.reshape(-1) for v in self.x.vars(x)]
        y1, y2 = [v.reshape(-1) for v in self.y.vars(y)]
        # y1 = x1 + tau K A^T x2 - K x1
        self.A.adjoint(x2, x1tmp)
        self.K(x1tmp, y1)
        _axpby(x1, y1, y1, 1.0, self.tau)
        self.K(x1, x1tmp)
        _axpby(y1, x1tmp, y1, 1.0, -1.0)
        # y2 = x2 - sigma H A x1 - H x2
        self.A(x1, x2tmp)
        self.H(x2tmp, y2)
        _axpby(x2, y2, y2, 1.0, -self.sigma)
        self.H(x2, x2tmp)
        _axpby(y2, x2tmp, y2, 1.0, -1.0)


class SemismoothNewtonSystemAdjoint(LinOp):
    """Adjoint of a SemismoothNewtonSystem M:
    [[I - K^T, -sigma A^T H^T], [tau A K^T, I - H^T]]
    (reference: semismooth_newton.py:221-248)."""

    def __init__(self, M):
        LinOp.__init__(self)
        self.M = M
        self.x = M.y
        self.y = M.x
        self.adjoint = M

    def _call_gpu(self, x, y=None, add=False):
        assert y is not None
        assert add is False
        M = self.M
        x1tmp, x2tmp = M._tmp(x)
        x1, x2 = [v.reshape(-1) for v in self.x.vars(x)]
        y1, y2 = [v.reshape(-1) for v in self.y.vars(y)]
        # y1 = x1 - sigma A^T H^T x2 - K^T x1, y2 = x2 + tau A K^T x1 - H^T x2
        M.H.adjoint(x2, x2tmp)
        M.K.adjoint(x1, x1tmp)
        M.A.adjoint(x2tmp, y1)
        _axpby(x1, y1, y1, 1.0, -M.sigma)
        _axpby(y1, x1tmp, y1, 1.0, -1.0)
        M.A(x1tmp, y2)
        _axpby(x2, y2, y2, 1.0, M.tau)
        _axpby(y2, x2tmp, y2, 1.0, -1.0)


def _flat_start(v, size, dtype):
    """A flat device copy of a start: a flat vector, or an (x, y) pair as
    PDHG.state returns it."""
    if isinstance(v, (tuple, list)):
        parts = [p.reshape(-1).to(dtype) if _hip.is_device_tensor(p)
                 else torch.as_tensor(np.asarray(p).ravel(), dtype=dtype).cuda() for p in v]
        v = torch.cat(parts)
    elif _hip.is_device_tensor(v):
        v = v.reshape(-1).to(dtype)
    else:
        v = torch.as_tensor(np.asarray(v).ravel(), dtype=dtype).cuda()
    if v.numel() != size:
        raise ValueError("SemismoothNewton: start has %d entries, the problem %d" % (v.numel(), size))
    return v


class SemismoothNewton(object):
    """SemismoothNewton(g, f, A).solve(...) with the reference's signature and
    result dict (reference: semismooth_newton.py:250-427)."""

    def __init__(self, g, f, A):
        self.g = g
        self.f = f
        self.linop = A
        self.xy = Variable((self.g.x.size,), (self.f.x.size,))
        self.itervars = {"xyk": None}
        self.constvars = {"tau": 1.0, "sigma": 1.0}

    def obj_primal(self, x, Ax):
        obj, infeas = self.g(x)
        obj2, infeas2 = self.f(Ax)
        return obj + obj2, infeas + infeas2

    def obj_dual(self, ATy, y):
        neg = torch.empty_like(ATy)
        _hip.call("scale", ATy.dtype, _hip.ptr(ATy), _hip.ptr(neg), ATy.numel(), None, -1.0, 0, _hip.stream())
        obj, infeas = self.g.conj(neg)
        obj2, infeas2 = self.f.conj(y)
        return -obj - obj2, infeas + infeas2

    def _split(self, v):
        return [t.reshape(-1) for t in self.xy.vars(v)]

    def res(self, xy, xyres, subdiff=False, norm=True):
        """xyres = F(xy); with subdiff the Jacobians of the two prox maps at
        this point go to the Newton system; returns 0.5 |F|^2 if norm
        (reference: semismooth_newton.py:269-291)."""
        c = self.constvars
        x, y = self._split(xy)
        xgrad, ygrad = self._split(xyres)
        # xgrad = x - prox_{tau g}(x - tau A^T y)
        self.linop.adjoint(y, xgrad)
        _axpby(x, xgrad, xgrad, 1.0, -c["tau"])
        K = self.gprox(xgrad, jacobian=subdiff)
        _axpby(x, xgrad, xgrad, 1.0, -1.0)
        # ygrad = y - prox_{sigma f*}(y + sigma A x)
        self.linop(x, ygrad)
        _axpby(y, ygrad, ygrad, 1.0, c["sigma"])
        H = self.fconjprox(ygrad, jacobian=subdiff)
        _axpby(y, ygrad, ygrad, 1.0, -1.0)
        if subdiff:
            if K is None or H is None:
                raise NotImplementedError("SemismoothNewton: a prox map of this problem has no Jacobian")
            self.M.K = K
            self.M.H = H
        if norm:
            return 0.5 * _dot(xyres, xyres)

    def iteration_step(self, _iter):
        """One damped Newton step (reference: semismooth_newton.py:293-331)."""
        i = self.itervars
        res_normsq = self.res(i["xyk"], i["xyres"], subdiff=True)
        logging.info("Start linsolve")
        i["cg_steps"] = CG_solve(self.M, i["xyres"], i["dk"], lbd=self.lbd,
                                 tmpvars=(i["cg_rk"], i["cg_rkp1"], i["xytmp"], i["cg_dk"]), tol=res_normsq)
        logging.info("Stop linsolve")
        self.M(i["dk"], i["xytmp"])
        p_gradf = _dot(i["xyres"], i["xytmp"])
        armijo_fun = lambda xy: self.res(xy, i["xyres"], subdiff=False)
        new_normsq, alpha = armijo(armijo_fun, i["xyk"], i["xykp1"], i["dk"], p_gradf, res_normsq)
        # trust-region-like update of the regularization (after "Distributed
        # Newton Methods for Deep Neural Networks", C.-C. Wang et al.)
        Md_normsq = 0.5 * _dot(i["xytmp"], i["xytmp"])
        rho = (new_normsq - res_normsq) / (alpha * p_gradf + alpha ** 2 * Md_normsq)
        if rho > 0.75:
            self.lbd *= self.lbd_drop
        elif rho < 0.25:
            self.lbd *= self.lbd_boost
        logging.debug("#{:6d}: alpha = {: 9.6g}, rho = {: 9.6g}, lbd = {: 9.6g}, normsq = {: 9.6g}"
                      .format(_iter, alpha, rho, self.lbd, res_normsq))
        _copy(i["xykp1"], i["xyk"])
        i["history"].append({"res_normsq": res_normsq, "alpha": alpha, "rho": rho, "lbd": self.lbd,
                             "cg_steps": i["cg_steps"]})

    def prepare_stepsizes(self, step_factor=1.0):
        """tau sigma |A|^2 = 1.25 x the truncated bound (reference:
        semismooth_newton.py:333-352)."""
        c = self.constvars
        logging.info("Estimating optimal step bound...")
        op_norm, itn = normest(self.linop)
        bnd = truncate(1.0 / op_norm ** 2, 3)  # < 1/|K|^2
        bnd *= 1.25  # boost
        fact = step_factor  # tau/sigma
        c["sigma"] = np.sqrt(bnd / fact)
        c["tau"] = bnd / c["sigma"]
        logging.info("Constant steps: %f (%f | %f)" % (bnd, c["sigma"], c["tau"]))
        self.gprox = self.g.prox(c["tau"])
        self.fconjprox = self.f.conj.prox(c["sigma"])
        self.M = SemismoothNewtonSystem(self.linop, c["tau"], c["sigma"])
        self.lbd = 1.0
        self.lbd_drop = 0.1
        self.lbd_boost = 5.0

    def solve(self, continue_at=None, granularity=50, term_relgap=1e-5, term_infeas=None, term_maxiter=int(5e2),
              precision="double"):
        """Damped Newton steps until the relative gap and the
        infeasibilities, checked every `granularity` steps, are below the
        thresholds (reference: semismooth_newton.py:354-424)."""
        _hip.require_device()
        i = self.itervars
        dtype = _hip.torch_dtype(precision)
        if continue_at is not None:
            i["xyk"] = _flat_start(continue_at, self.xy.size, dtype)
        elif i.get("xyk") is None or not _hip.is_device_tensor(i["xyk"]):
            i["xyk"] = torch.zeros(self.xy.size, dtype=dtype, device="cuda")
        else:
            i["xyk"] = i["xyk"].to(dtype)
        for name in ["xykp1", "xyres", "dk", "xytmp", "cg_dk", "cg_rk", "cg_rkp1"]:
            i[name] = i["xyk"].clone()
        i["history"] = []
        xk, yk = self._split(i["xyk"])
        xkp1, ykp1 = self._split(i["xykp1"])

        self.prepare_stepsizes()

        if term_infeas is None:
            term_infeas = term_relgap
        obj_p = obj_d = infeas_p = infeas_d = relgap = 0.
        self.checks = []  # one record per gap check

        logging.info("Solving (steps<%d)..." % term_maxiter)
        with GracefulInterruptHandler() as interrupt_hdl:
            _iter = 0
            while _iter < term_maxiter:
                self.iteration_step(_iter)
                _iter += 1
                if interrupt_hdl.interrupted or _iter % granularity == 0:
                    if interrupt_hdl.interrupted:
                        print("Interrupt (SIGINT) at iter=%d" % _iter)
                    self.linop(xk, ykp1)
                    self.linop.adjoint(yk, xkp1)
                    obj_p, infeas_p = self.obj_primal(xk, ykp1)
                    obj_d, infeas_d = self.obj_dual(xkp1, yk)
                    # relative primal-dual gap, as the reference writes it here
                    relgap = (obj_p - obj_d) / max(np.spacing(1), obj_d)
                    self.checks.append({"iter": _iter, "objp": obj_p, "objd": obj_d, "infeasp": infeas_p,
                                        "infeasd": infeas_d, "relgap": relgap})
                    logging.info("#{:6d}: objp = {: 9.6g} ({: 9.6g}), objd = {: 9.6g} ({: 9.6g}), "
                                 "gap = {: 9.6g}, relgap = {: 9.6g} ".format(
                                     _iter, obj_p, infeas_p, obj_d, infeas_d, obj_p - obj_d, relgap))
                    if np.abs(relgap) < term_relgap and max(infeas_p, infeas_d) < term_infeas:
                        break
                    if interrupt_hdl.interrupted:
                        break
        self.iterations = _iter
        return {"objp": obj_p, "objd": obj_d, "infeasp": infeas_p, "infeasd": infeas_d, "relgap": relgap}

    @property
    def state(self):
        """(x, y) of the current iterate on the host, like PDHG.state."""
        return tuple(t.cpu().numpy() for t in self._split(self.itervars["xyk"]))


def _norm(v):
    """Euclidean norm as numpy's norm forms it for a vector: sqrt(<v, v>)."""
    d = _dot(v, v)
    return math.sqrt(d) if d >= 0 else float("nan")


def _scalar_search_armijo(phi, phi0, derphi0, c1=1e-4, alpha0=1.0, amin=0.0):
    """Backtracking with quadratic, then cubic interpolation (Nocedal & Wright,
    Numerical Optimization, 1999, pp. 56-57), as scipy's line search of the
    same name runs it (scipy 1.15 optimize/_linesearch.py:684-739).  Returns
    (alpha or None, phi there)."""
    phi_a0 = phi(alpha0)
    if phi_a0 <= phi0 + c1 * alpha0 * derphi0:
        return alpha0, phi_a0
    # minimizer of the quadratic through phi0, derphi0 and phi_a0
    alpha1 = -(derphi0) * alpha0 ** 2 / 2.0 / (phi_a0 - phi0 - derphi0 * alpha0)
    phi_a1 = phi(alpha1)
    if phi_a1 <= phi0 + c1 * alpha1 * derphi0:
        return alpha1, phi_a1
    # then the minimizers of cubics through phi0, derphi0 and the last two values
    while alpha1 > amin:
        factor = alpha0 ** 2 * alpha1 ** 2 * (alpha1 - alpha0)
        a = alpha0 ** 2 * (phi_a1 - phi0 - derphi0 * alpha1) - alpha1 ** 2 * (phi_a0 - phi0 - derphi0 * alpha0)
        a = a / factor
        b = -alpha0 ** 3 * (phi_a1 - phi0 - derphi0 * alpha1) + alpha1 ** 3 * (phi_a0 - phi0 - derphi0 * alpha0)
        b = b / factor
        alpha2 = (-b + math.sqrt(abs(b ** 2 - 3 * a * derphi0))) / (3.0 * a)
        phi_a2 = phi(alpha2)
        if phi_a2 <= phi0 + c1 * alpha2 * derphi0:
            return alpha2, phi_a2
        if (alpha1 - alpha2) > alpha1 / 2.0 or (1 - alpha2 / alpha1) < 0.96:
            alpha2 = alpha1 / 2.0
        alpha0, alpha1 = alpha1, alpha2
        phi_a0, phi_a1 = phi_a1, phi_a2
    return None, phi_a1


class _LowRankInverse(object):
    """G = g I + sum_n c_n d_n^T on device vectors (scipy 1.15
    optimize/_nonlin.py:614-700, LowRankMatrix without the dense collapse,
    which only a rank above the vector length would trigger)."""

    def __init__(self, g):
        self.g = g
        self.cs, self.ds = [], []

    def matvec(self, v, out):
        """out = G v"""
        _axpby(v, v, out, self.g, 0.0)
        for c, d in zip(self.cs, self.ds):
            _axpby(c, out, out, _dot(d, v), 1.0)

    def rmatvec(self, v, out):
        """out = G^T v"""
        _axpby(v, v, out, self.g, 0.0)
        for c, d in zip(self.cs, self.ds):
            _axpby(d, out, out, _dot(c, v), 1.0)


def broyden1(F, x0, iters):
    """`iters` steps of Broyden's first ("good") method with Armijo line
    search from x0, as scipy.optimize.broyden1(F, x0, iter=iters) takes them
    with its defaults (scipy 1.15 optimize/_nonlin.py: nonlin_solve :122-270,
    _nonlin_line_search :272-315, TerminationCondition :317-375, GenericBroyden
    :588-612, BroydenFirst :842-905): the inverse Jacobian approximation
    starts at -alpha I with alpha = 0.5 max(|x0|, 1)/|F(x0)| and takes one
    rank-one term per step.  F(x, out) writes F(x) into out.  Returns x on
    the device.

    A non-finite step restarts the approximation at -alpha I, as scipy's does;
    here the test is on <r, r> (one FP64 reduction), which differs from
    scipy's entrywise test only for entries beyond 1e154, and d = v/<df, v>
    is formed as a product with the reciprocal (one rounding more)."""
    x = x0.clone()
    Fx = torch.empty_like(x)
    F(x, Fx)

    last_x, last_f = x.clone(), Fx.clone()
    nf = _norm(Fx)
    alpha = 0.5 * max(_norm(x), 1) / nf if nf else 1.0  # set once; a restart keeps it
    G = _LowRankInverse(-alpha)
    dx = torch.empty_like(x)
    it = 0
    while True:
        it += 1  # TerminationCondition.check with iter given: stop on F = 0 or after `iters` steps
        if _dot(Fx, Fx) == 0 or it > iters:
            break
        G.matvec(Fx, dx)
        if not math.isfinite(_dot(dx, dx)):
            G = _LowRankInverse(-alpha)
            G.matvec(Fx, dx)
        _axpby(dx, dx, dx, -1.0, 0.0)  # dx = -G F(x)
        if _norm(dx) == 0:
            raise ValueError("broyden1: the inverse Jacobian estimate maps F(x) to zero")
        # line search on phi(s) = |F(x + s dx)|^2 (scipy _nonlin_line_search, 'armijo')
        store = {"s": 0.0, "phi": _norm(Fx) ** 2, "F": Fx}

        def phi(sv):
            if sv == store["s"]:
                return store["phi"]
            xt = torch.empty_like(x)
            _axpby(x, dx, xt, 1.0, sv)
            v = torch.empty_like(x)
            F(xt, v)
            d = _dot(v, v)
            p = math.sqrt(d) ** 2 if math.isfinite(d) else math.inf
            store.update(s=sv, phi=p, F=v)
            return p

        sv, _ = _scalar_search_armijo(phi, store["phi"], -store["phi"], amin=1e-2)
        if sv is None:
            sv = 1.0
        _axpby(x, dx, x, 1.0, sv)
        if sv == store["s"]:
            Fx = store["F"]
        else:
            Fx = torch.empty_like(x)
            F(x, Fx)
        # Broyden update of the inverse: H+ = H + (dx - H df) dx^T H / (dx^T H df)
        df = torch.empty_like(x)
        _axpby(Fx, last_f, df, 1.0, -1.0)
        ddx = torch.empty_like(x)
        _axpby(x, last_x, ddx, 1.0, -1.0)
        v = torch.empty_like(x)
        G.rmatvec(ddx, v)
        c = torch.empty_like(x)
        G.matvec(df, c)
        _axpby(ddx, c, c, 1.0, -1.0)
        d = v
        _axpby(v, v, d, 1.0 / _dot(df, v), 0.0)
        G.cs.append(c)
        G.ds.append(d)
        last_f = Fx
        last_x = x.clone()
    return x


class SemismoothQuasinewton(SemismoothNewton):
    """The fixed-point residual of PDHG solved by Broyden's first method in
    rounds of `granularity` steps, each round from a fresh Jacobian
    approximation and followed by a gap check (reference:
    semismooth_newton.py:429-500, where each round is one call of
    scipy.optimize.broyden1)."""

    def solve(self, continue_at=None, granularity=10, term_relgap=1e-5, term_infeas=None, term_maxiter=int(1e2),
              precision="double"):
        _hip.require_device()
        i = self.itervars
        dtype = _hip.torch_dtype(precision)
        if continue_at is not None:
            i["xyk"] = _flat_start(continue_at, self.xy.size, dtype)
        elif i.get("xyk") is None or not _hip.is_device_tensor(i["xyk"]):
            i["xyk"] = torch.zeros(self.xy.size, dtype=dtype, device="cuda")
        else:
            i["xyk"] = i["xyk"].to(dtype)
        i["xykp1"] = i["xyk"].clone()
        i["xykres"] = i["xyk"].clone()
        xk, yk = self._split(i["xyk"])
        xkp1, ykp1 = self._split(i["xykp1"])

        self.prepare_stepsizes()

        def f(x, out):
            self.res(x, out, norm=False)

        if term_infeas is None:
            term_infeas = term_relgap
        obj_p = obj_d = infeas_p = infeas_d = relgap = 0.
        self.checks = []
        logging.info("Solving (steps<%d)..." % term_maxiter)
        with GracefulInterruptHandler() as interrupt_hdl:
            _iter = 0
            while _iter < term_maxiter:
                _copy(i["xyk"], i["xykp1"])
                _copy(broyden1(f, i["xykp1"], granularity), i["xyk"])
                _iter += granularity
                if interrupt_hdl.interrupted or _iter % granularity == 0:
                    if interrupt_hdl.interrupted:
                        print("Interrupt (SIGINT) at iter=%d" % _iter)
                    self.linop(xk, ykp1)
                    self.linop.adjoint(yk, xkp1)
                    obj_p, infeas_p = self.obj_primal(xk, ykp1)
                    obj_d, infeas_d = self.obj_dual(xkp1, yk)
                    relgap = (obj_p - obj_d) / max(np.spacing(1), obj_d)
                    self.checks.append({"iter": _iter, "objp": obj_p, "objd": obj_d, "infeasp": infeas_p,
                                        "infeasd": infeas_d, "relgap": relgap})
                    logging.info("#{:6d}: objp = {: 9.6g} ({: 9.6g}), objd = {: 9.6g} ({: 9.6g}), "
                                 "gap = {: 9.6g}, relgap = {: 9.6g} ".format(
                                     _iter, obj_p, infeas_p, obj_d, infeas_d, obj_p - obj_d, relgap))
                    if np.abs(relgap) < term_relgap and max(infeas_p, infeas_d) < term_infeas:
                        break
                    if interrupt_hdl.interrupted:
                        break
        self.iterations = _iter
        return {"objp": obj_p, "objd": obj_d, "infeasp": infeas_p, "infeasd": infeas_d, "relgap": relgap}


__all__ = ["armijo", "SymRegOp", "CG_solve", "SemismoothNewtonSystem", "SemismoothNewtonSystemAdjoint",
           "SemismoothNewton", "SemismoothQuasinewton", "broyden1"]
