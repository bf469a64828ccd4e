"""Indicator functionals (reference: opymize/functionals/indicators.py)."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.operators.affine import ConstOp, ConstrainOp
from opymize_amd.operators.proj import PosProj, NegProj, ClipOp, SimplexProj


class IndicatorFct(Functional):
    """F(x) = c2 if x == c1 else inf (reference: indicators.py:7-28)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, c1=0, c2=0, conj=None):
        Functional.__init__(self)
        self.c1 = c1
        self.c2 = c2
        self.x = Variable(N)
        if conj is None:
            from opymize_amd.functionals.affine import AffineFct
            self.conj = AffineFct(N, b=-c2, c=c1, conj=self)
        else:
            self.conj = conj
        self._prox = ConstOp(N, self.x.new() + c1)
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        if isinstance(self.c1, np.ndarray):
            prm.c = vptr(self._const.get("c1", self.c1.ravel(), x.dtype))
        else:
            prm.cs = float(self.c1)
        fn_eval(_hip.FN_INDICATOR, x, 1, out, prm)

    def _gpu_finish(self, v):
        return float(self.c2), float(v[1])

    def prox(self, tau):
        return self._prox


class ConstrainFct(Functional):
    """F(x) = 0 if x[mask,:] == c[mask,:] else inf (reference: indicators.py:30-57)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, mask, c, conj=None):
        Functional.__init__(self)
        self.c = np.asarray(c)
        self.x = Variable(self.c.shape)
        self.mask = np.asarray(mask).astype(bool)
        if conj is None:
            from opymize_amd.functionals.affine import MaskedAffineFct
            self.conj = MaskedAffineFct(mask, c, conj=self)
        else:
            self.conj = conj
        self._prox = ConstrainOp(self.mask, self.c)
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.c = vptr(self._const.get("c", self.c.ravel(), x.dtype))
        prm.mask = vptr(self._const.get_u8("mask", self.mask))
        fn_eval(_hip.FN_INDICATOR, x, self.c.shape[1], out, prm)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class PositivityFct(Functional):
    """F(x) = 0 if x >= 0 else inf (reference: indicators.py:59-77)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, conj=None):
        Functional.__init__(self)
        self.x = Variable(N)
        self.conj = NegativityFct(N, conj=self) if conj is None else conj
        self._prox = PosProj(N)

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_POS, x, 1, out)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class NegativityFct(Functional):
    """F(x) = 0 if x <= 0 else inf (reference: indicators.py:79-98)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, conj=None):
        Functional.__init__(self)
        self.x = Variable(N)
        self.conj = PositivityFct(N, conj=self) if conj is None else conj
        self._prox = NegProj(N)

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_NEG, x, 1, out)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class BoxFct(Functional):
    """F(x) = 0 if lo <= x <= hi entrywise, else inf (no reference
    counterpart: the range of the lifted variable of the multi-label lifting
    model).  Its conjugate is the support function BoxSupp, finite everywhere,
    so a dual objective with a box-constrained primal carries no
    infeasibility; the prox is ClipOp."""
    device_step = True  # prox is independent of the step

    def __init__(self, N, lo=0.0, hi=1.0, conj=None):
        Functional.__init__(self)
        if not lo <= hi:
            raise ValueError("BoxFct needs lo <= hi")
        self.x = Variable(N)
        self.lo, self.hi = float(lo), float(hi)
        self.conj = BoxSupp(N, lo, hi, conj=self) if conj is None else conj
        self._prox = ClipOp(N, self.lo, self.hi)

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.cs, prm.lbd = self.lo, self.hi
        fn_eval(_hip.FN_BOX_IND, x, 1, out, prm)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class BoxSupp(Functional):
    """F(z) = sum_i max(lo z_i, hi z_i), the support function of [lo, hi]^N
    (conjugate of BoxFct)."""

    def __init__(self, N, lo=0.0, hi=1.0, conj=None):
        Functional.__init__(self)
        self.x = Variable(N)
        self.lo, self.hi = float(lo), float(hi)
        self.conj = BoxFct(N, lo, hi, conj=self) if conj is None else conj

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.cs, prm.lbd = self.lo, self.hi
        fn_eval(_hip.FN_BOX_SUPP, x, 1, out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), 0.0


class SimplexFct(Functional):
    """F(x) = 0 if every point's K values lie in the probability simplex
    {v >= 0, sum v = 1}, else inf; infeasibility max(|sum v - 1|, max(0, -v))
    (no reference counterpart: the piece weights of the multi-label lifting
    model).  Conjugate SimplexSupp (finite everywhere); prox SimplexProj."""
    device_step = True  # prox is independent of the step

    def __init__(self, N, K, conj=None):
        Functional.__init__(self)
        self.N, self.K = int(N), int(K)
        self.x = Variable((self.N, self.K))
        self.conj = SimplexSupp(N, K, conj=self) if conj is None else conj
        self._prox = SimplexProj(N, K)
        self.precond_block = self.K  # the projection couples each point's K entries

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_SIMPLEX_IND, x, self.K, out)

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox


class SimplexSupp(Functional):
    """F(z) = sum over points of max_k z_k, the support function of the
    simplex per point (conjugate of SimplexFct)."""

    def __init__(self, N, K, conj=None):
        Functional.__init__(self)
        self.N, self.K = int(N), int(K)
        self.x = Variable((self.N, self.K))
        self.conj = SimplexFct(N, K, conj=self) if conj is None else conj

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_SIMPLEX_SUPP, x, self.K, out)

    def _gpu_finish(self, v):
        return float(v[0]), 0.0
