"""Epigraph functionals of quadratics (reference: opymize/functionals/epigraph.py:129-274)."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.operators.proj import QuadEpiProj


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

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.qa = self.alph
        prm.lbd = self.lbd
        fn_eval(_hip.FN_TRUNCQUADEPI, x, self.M + 1, out, prm)

    def _gpu_finish(self, v):
        return 0.0, float(v[1]) + float(v[2])

    def prox(self, tau):
        return self._prox
