"""Affine functionals (reference: opymize/functionals/affine.py)."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.linear.scale import IdentityOp
from opymize_amd.operators.affine import ShiftScaleOp


class AffineFct(Functional):
    """F(x) = b + sum_i c_i x_i (reference: affine.py:7-40).

    Like the reference, the conjugate is IndicatorFct(c1=c, c2=b)
    (reference affine.py:15), i.e. its value carries +b."""

    def __init__(self, N, b=0, c=1, conj=None):
        Functional.__init__(self)
        self.b = b
        self.c = c
        self.x = Variable(N)
        if conj is None:
            from opymize_amd.functionals.indicators import IndicatorFct
            self.conj = IndicatorFct(N, c1=c, c2=b, conj=self)
        else:
            self.conj = conj
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        if isinstance(self.c, np.ndarray):
            prm.c = vptr(self._const.get("c", self.c.ravel(), x.dtype))
        else:
            prm.cs = float(self.c)
        fn_eval(_hip.FN_DOT, x, 1, out, prm)

    def _gpu_finish(self, v):
        return float(self.b) + float(v[0]), 0.0

    def prox(self, tau):
        # x - tau*c  (reference: ShiftScaleOp(N, c, 1, -tau))
        return ShiftScaleOp(self.x.size, self.c, 1, -tau)


class ConstFct(AffineFct):
    """F(x) = const (reference: affine.py:42-45)."""

    def __init__(self, N, const=0, conj=None):
        AffineFct.__init__(self, N, b=const, c=0)


class ZeroFct(ConstFct):
    """F(x) = 0 (reference: affine.py:47-54)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, conj=None):
        ConstFct.__init__(self, N, const=0)

    def prox(self, tau):
        return IdentityOp(self.x.size)


class MaskedAffineFct(Functional):
    """F(x) = sum(c[mask,:]*x[mask,:]) + indicator(x[~mask,:] == 0)
    (reference: affine.py:56-88).  Its conjugate is ConstrainFct(mask, c)."""

    def __init__(self, mask, c, conj=None):
        Functional.__init__(self)
        self.c = np.asarray(c)
        self.x = Variable(self.c.shape)
        self.mask = np.asarray(mask).astype(bool)
        self.nmask = ~self.mask
        if conj is None:
            from opymize_amd.functionals.indicators import ConstrainFct
            self.conj = ConstrainFct(mask, c, conj=self)
        else:
            self.conj = conj
        scale = np.zeros(self.c.shape)
        scale[self.mask, :] = 1.0
        self._scale = scale.ravel()
        self._const = DeviceConstants()

    def _gpu_eval(self, x, out):
        prm = _hip.FnParams()
        prm.c = vptr(self._const.get("c", self.c.ravel(), x.dtype))
        prm.mask = vptr(self._const.get_u8("mask", self.mask))
        fn_eval(_hip.FN_MASKED_AFFINE, x, self.c.shape[1], out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), (0.0 if self.mask.all() else float(v[1]))

    def prox(self, tau):
        # scale * (x - tau*c)  (reference: ShiftScaleOp(N, c, scale, -tau))
        return ShiftScaleOp(self.x.size, self.c.ravel(), self._scale, -tau)
