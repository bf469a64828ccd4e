"""Quadratic data terms: SSD (reference: opymize/functionals/quad.py) and the
TV-L1 data term L1Data (no counterpart in the reference)."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.functionals._base import fn_eval, vptr
from opymize_amd.operators.affine import SSDProxOp, L1DataProxOp


class SSD(Functional):
    """G(x) = 0.5*<x - f, x - f>_vol + shift over unmasked points
    (reference: quad.py:7-55).  The conjugate is the SSD with data -vol*f,
    volume 1/vol and shift -0.5*sum(vol f^2) - shift."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, data, vol=None, shift=0, mask=None, conj=None):
        Functional.__init__(self)
        data = np.asarray(data)
        self.x = Variable(data.shape)
        self.f = np.atleast_2d(data).astype(np.float64)
        self.shift = float(shift)
        self.vol = np.ones(self.f.shape[1]) if vol is None else np.asarray(vol, dtype=np.float64).ravel()
        self.mask = np.ones(self.f.shape[0], dtype=bool) if mask is None else np.asarray(mask).astype(bool)
        if conj is None:
            cj_vol = 1.0 / self.vol
            cj_data = np.zeros_like(self.f)
            cj_data[self.mask, :] = self.f[self.mask, :] * (-self.vol)[None, :]
            cj_shift = -0.5 * np.einsum("ik,k->", cj_data ** 2, cj_vol)
            cj_shift -= self.shift
            self.conj = SSD(cj_data, shift=cj_shift, vol=cj_vol, mask=mask, conj=self)
        else:
            self.conj = conj
        self._const = DeviceConstants()

    def _dev(self, dtype):
        f = self._const.get("f", self.f.ravel(), dtype)
        vol = self._const.get("vol", self.vol, _hip.torch_dtype("double"))
        mask = None if self.mask.all() else self._const.get_u8("mask", self.mask)
        return f, vol, mask

    def _gpu_eval(self, x, out):
        f, vol, mask = self._dev(x.dtype)
        prm = _hip.FnParams()
        prm.f, prm.vol, prm.mask = vptr(f), vptr(vol), vptr(mask)
        fn_eval(_hip.FN_SSD, x, self.f.shape[1], out, prm)

    def _gpu_finish(self, v):
        return float(v[0]) + self.shift, 0.0

    def prox(self, tau):
        return SSDProxOp(self, tau)


class L1Data(Functional):
    """G(x) = sum_ik lbd_k |x_ik - f_ik|, the data term of TV-L1 models.
    The reference has no such functional (its L1Norms has no shift and no
    prox).  prox = soft threshold about f; conjugate: L1DataConj."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, data, lbd=None, conj=None):
        Functional.__init__(self)
        data = np.asarray(data)
        self.x = Variable(data.shape)
        self.f = np.atleast_2d(data).astype(np.float64)
        C = self.f.shape[1]
        self.lbd = np.ones(C) if lbd is None else np.broadcast_to(np.asarray(lbd, dtype=np.float64), (C,)).copy()
        self._const = DeviceConstants()
        self.conj = L1DataConj(self) if conj is None else conj

    def _dev(self, dtype):
        return (self._const.get("f", self.f.ravel(), dtype),
                self._const.get("lbd", self.lbd, _hip.torch_dtype("double")))

    def _gpu_eval(self, x, out):
        f, lbd = self._dev(x.dtype)
        prm = _hip.FnParams()
        prm.f, prm.vol = vptr(f), vptr(lbd)
        fn_eval(_hip.FN_L1DATA, x, self.f.shape[1], out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), 0.0

    def prox(self, tau):
        return L1DataProxOp(self, tau, conj=False)


class L1DataConj(Functional):
    """G*(z) = sum_ik z_ik f_ik + indicator(|z_ik| <= lbd_k); infeasibility
    is the largest violation max(0, |z| - lbd)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, primal):
        Functional.__init__(self)
        self.x = primal.x
        self.f = primal.f
        self.lbd = primal.lbd
        self.conj = primal
        self._primal = primal

    def _dev(self, dtype):
        return self._primal._dev(dtype)

    def _gpu_eval(self, x, out):
        f, lbd = self._dev(x.dtype)
        prm = _hip.FnParams()
        prm.f, prm.vol = vptr(f), vptr(lbd)
        fn_eval(_hip.FN_L1DATA_CONJ, x, self.f.shape[1], out, prm)

    def _gpu_finish(self, v):
        return float(v[0]), float(v[1])

    def prox(self, tau):
        return L1DataProxOp(self._primal, tau, conj=True)
