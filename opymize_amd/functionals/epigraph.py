"""Epigraph functionals: polyhedral (reference: opymize/functionals/epigraph.py:7-127)
and of quadratics (reference: opymize/functionals/epigraph.py:129-274)."""
import numpy as np
import torch

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.operators.proj import QuadEpiProj, EpigraphProj, epigraph_Ab


def _epigraph_arrays(I, J, v, b):
    I = np.asarray(I, dtype=bool)
    J = np.asarray(J)
    v = np.asarray(v, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert I.shape == b.shape and I.shape[1] == v.shape[0]
    return I, J, v, b


def _epi_params(const, I, J, v, b):
    prm = _hip.FnParams()
    prm.ei = vptr(const.get_u8("I", I))
    prm.ej = vptr(const.get("J", J.astype(np.int32), torch.int32))
    prm.ev = vptr(const.get("v", v, torch.float64))
    prm.eb = vptr(const.get("b", b, torch.float64))
    prm.nfuns, prm.npoints = I.shape
    prm.nsub = J.shape[1]
    return prm


class EpigraphInd(Functional):
    """F(x) = 0 if <v[l], x[j,i,:-1]> - b[i,l] <= x[j,i,-1] for all j, i and
    l in J[j] with I[i,l], else infinity; value 0, infeasibility the largest
    violation (reference: epigraph.py:7-53).  The prox is EpigraphProj."""
    device_step = True  # prox is independent of the step

    def __init__(self, I, If, J, v, b, conj=None):
        Functional.__init__(self)
        I, J, v, b = _epigraph_arrays(I, J, v, b)
        nfuns, npoints = I.shape
        nregions, nsub = J.shape
        ndim = v.shape[1]
        self.I, self.J, self.v = I, J, v
        self.x = Variable((nregions, nfuns, ndim + 1))
        self.conj = EpigraphSupp(I, If, J, v, b, conj=self) if conj is None else conj
        self.A = self.conj.A
        self.b = self.conj.b  # right-hand side of the constraint system, as in the reference
        self._fb = b
        self._prox = EpigraphProj(I, J, v, b, Ab=(self.A, self.b))
        self.precond_block = self.v.shape[1] + 1  # one QP per (region, function) block
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _epi_params(self._const, self.I, self.J, self.v, self._fb)
        fn_eval(_hip.FN_EPIGRAPH_IND, x, self.v.shape[1] + 1, out, prm)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


EpigraphFct = EpigraphInd


class EpigraphSupp(Functional):
    """F(x) = max sum_ji <y, x[j,i]> over y[j,i] in epi(f[i]_j*), i.e. the
    perspective xt * f[i]_j(x[:-1]/xt) with xt = |x[-1]| evaluated through the
    affine pieces of f[i]_j given by the faces If[i][j] (index triples of
    J[j]); infeasibility max(0, -checks[j] x[j,i]) with checks[j] the inverse
    transpose of the rows (v[J[j,k]], -1), k <= ndim (reference:
    epigraph.py:55-127)."""

    def __init__(self, I, If, J, v, b, conj=None):
        Functional.__init__(self)
        I, J, v, b = _epigraph_arrays(I, J, v, b)
        nfuns, npoints = I.shape
        nregions, nsub = J.shape
        ndim = v.shape[1]
        self.I, self.J, self.v, self._fb = I, J, v, b
        self.x = Variable((nregions, nfuns, ndim + 1))
        self.A, self.b = epigraph_Ab(I, J, v, b)

        # affine pieces g.x + c through the ndim+1 points of each face
        self.eqns = []
        for j in range(nregions):
            for i in range(nfuns):
                faces = np.asarray(If[i][j], dtype=np.int64).reshape(-1, ndim + 1)
                pts = v[J[j]][faces]             # (nfaces, ndim+1, ndim)
                vals = b[i, J[j]][faces]         # (nfaces, ndim+1)
                eq = np.zeros((faces.shape[0], ndim + 1))
                if faces.shape[0]:
                    eq[:, :-1] = np.linalg.solve(pts[:, 1:] - pts[:, :1], (vals[:, 1:] - vals[:, :1])[..., None])[..., 0]
                    eq[:, -1] = vals[:, 0] - np.einsum("fd,fd->f", eq[:, :-1], pts[:, 0])
                self.eqns.append(eq)
        rows = -np.ones((nregions, ndim + 1, ndim + 1))
        rows[:, :, :-1] = v[J[:, :ndim + 1]]
        self.checks = np.linalg.inv(rows).transpose(0, 2, 1)
        self._eq = np.concatenate(self.eqns).reshape(-1) if self.eqns else np.zeros(0)
        self._eq_off = np.concatenate([[0], np.cumsum([e.shape[0] for e in self.eqns])]).astype(np.int64)
        self.conj = EpigraphInd(I, If, J, v, b, conj=self) if conj is None else conj
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.eq = vptr(self._const.get("eq", self._eq if self._eq.size else np.zeros(1), torch.float64))
        prm.eq_off = vptr(self._const.get("eq_off", self._eq_off, torch.int64))
        prm.checks = vptr(self._const.get("checks", self.checks.reshape(-1), torch.float64))
        prm.nfuns = self.I.shape[0]
        fn_eval(_hip.FN_EPIGRAPH_SUPP, x, self.v.shape[1] + 1, out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), float(v[1])


EpigraphSupportFct = EpigraphSupp


def quad_dual_coefficients(a, b, c):
    """f_i(x) = 0.5 a |x|^2 + <b_i, x> + c_i  =>  f_i*(x) = 0.5/a |x|^2 + <-b_i/a, x> + (0.5/a |b_i|^2 - c_i)
    (reference: epigraph.py:129-133)."""
    return (1.0 / a, -b / a, 0.5 / a * (b ** 2).sum(axis=-1) - c)


class _QuadBase(Functional):
    def __init__(self, N, M, a, b, c):
        Functional.__init__(self)
        assert a > 0
        self.N, self.M = int(N), int(M)
        self.x = Variable((self.N, self.M + 1))
        self.a = float(a)
        self.b = np.zeros((N, M)) if b is None else np.asarray(b, dtype=np.float64).reshape(N, M)
        self.c = np.zeros((N,)) if c is None else np.asarray(c, dtype=np.float64).reshape(N)
        self._const = DeviceConstants()

    def _params(self):
        dd = _hip.torch_dtype("double")
        prm = _hip.FnParams()
        prm.qa = self.a
        prm.qb = vptr(self._const.get("b", self.b.ravel(), dd))
        prm.qc = vptr(self._const.get("c", self.c, dd))
        return prm


class QuadEpiSupp(_QuadBase):
    """sum_i -x[i,-1] f_i(-x[i,:-1]/x[i,-1]) if x[i,-1] < 0, else inf
    (reference: epigraph.py:135-167)."""

    def __init__(self, N, M, a=1.0, b=None, c=None, conj=None):
        _QuadBase.__init__(self, N, M, a, b, c)
        if conj is None:
            da, db, dc = quad_dual_coefficients(self.a, self.b, self.c)
            self.conj = QuadEpiInd(N, M, a=da, b=db, c=dc, conj=self)
        else:
            self.conj = conj

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_QUADEPI_SUPP, x, self.M + 1, out, self._params())

    def _gpu_finish(self, v):
        return float(v[0]), float(v[1])


QuadSupport = QuadEpiSupp


class QuadEpiInd(_QuadBase):
    """sum_i indicator(f_i(x[i,:-1]) <= x[i,-1]) (reference: epigraph.py:169-201)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, M, a=1.0, b=None, c=None, conj=None):
        _QuadBase.__init__(self, N, M, a, b, c)
        if conj is None:
            da, db, dc = quad_dual_coefficients(self.a, self.b, self.c)
            self.conj = QuadEpiSupp(N, M, a=da, b=db, c=dc, conj=self)
        else:
            self.conj = conj
        self._prox = QuadEpiProj(N, M, alph=self.a, b=b, c=c)
        self.precond_block = self.M + 1  # the projection couples each point's M+1 entries

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_QUADEPI_IND, x, self.M + 1, out, self._params())

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class HuberPerspective(Functional):
    """Perspective of the Huber function lbd*(0.5/alph |x|^2 or |x| - alph/2)
    (reference: epigraph.py:203-236)."""

    def __init__(self, N, M, lbd=1.0, alph=1.0, conj=None):
        Functional.__init__(self)
        assert lbd > 0
        assert alph > 0
        self.N, self.M = int(N), int(M)
        self.x = Variable((self.N, self.M + 1))
        self.lbd = float(lbd)
        self.alph = float(alph)
        if conj is None:
            self.conj = TruncQuadEpiInd(N, M, lbd=self.lbd, alph=self.alph / self.lbd, conj=self)
        else:
            self.conj = conj

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.qa = self.alph
        prm.lbd = self.lbd
        fn_eval(_hip.FN_HUBER, x, self.M + 1, out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), float(v[1])


class TruncQuadEpiInd(Functional):
    """sum_i indicator(|x[i,:-1]| <= lbd) + indicator(0.5 alph |x[i,:-1]|^2 <= x[i,-1])
    (reference: epigraph.py:238-274)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, M, lbd=1.0, alph=1.0, conj=None):
        Functional.__init__(self)
        assert lbd > 0
        assert alph > 0
        self.N, self.M = int(N), int(M)
        self.x = Variable((self.N, self.M + 1))
        self.lbd = float(lbd)
        self.alph = float(alph)
        if conj is None:
            self.conj = HuberPerspective(N, M, lbd=self.lbd, alph=self.alph * self.lbd, conj=self)
        else:
            self.conj = conj
        self._prox = QuadEpiProj(N, M, lbd=self.lbd, alph=self.alph)
        self.precond_block = self.M + 1

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.qa = self.alph
        prm.lbd = self.lbd
        fn_eval(_hip.FN_TRUNCQUADEPI, x, self.M + 1, out, prm)

    def _gpu_finish(self, v):
        return 0.0, float(v[1]) + float(v[2])

    def prox(self, tau):
        return self._prox
