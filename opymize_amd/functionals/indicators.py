"""Indicator functionals (reference: opymize/functionals/indicators.py)."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.operators.affine import ConstOp, ConstrainOp
from opymize_amd.operators.proj import PosProj, NegProj


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
