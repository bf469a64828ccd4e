"""Affine prox maps (reference: opymize/operators/affine.py) and the SSD /
L1-data prox maps whose step size may live on the device."""
import numpy as np
import torch

from opymize_amd import Operator, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.linear.scale import ScaleOp, ZeroOp


def _out(x, y):
    return x if y is None else y


class StepArg(object):
    """A step size for a prox kernel: float, host array, device array, or a
    device scalar given as (float64 tensor, index)."""

    def __init__(self, tau):
        self.tau = tau
        self._const = DeviceConstants()

    def step(self, dtype):
        t = self.tau
        if isinstance(t, np.ndarray):
            if t.size == 1:
                return _hip.make_step(float(t.ravel()[0]), dtype)
            return _hip.make_step(self._const.get("tau", t.ravel(), dtype), dtype)
        return _hip.make_step(t, dtype)

    def host_value(self):
        """Scalar value on the host (only for host scalars)."""
        if isinstance(self.tau, (float, int, np.floating, np.integer)):
            return float(self.tau)
        return None


class ConstOp(Operator):
    """T(x) = const (reference: affine.py:15-46)."""

    def __init__(self, N, const):
        Operator.__init__(self)
        self.x = Variable(N)
        self.y = Variable(np.asarray(const).size)
        self.const = np.asarray(const)
        self._jacobian = ZeroOp(self.x.size, self.y.size)
        self._const = DeviceConstants()

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        y = _out(x, y)
        if y.numel() != self.const.size:
            raise ValueError("ConstOp: output has %d entries, the constant %d" % (y.numel(), self.const.size))
        c = self._const.get("c", self.const.ravel(), y.dtype)
        _hip.call("const", y.dtype, _hip.ptr(y), y.numel(), _hip.ptr(c), int(add), _hip.stream())
        if jacobian:
            return self._jacobian


class ConstrainOp(Operator):
    """T(x)[mask,:] = const[mask,:], identity elsewhere (reference: affine.py:48-83)."""

    def __init__(self, mask, const):
        Operator.__init__(self)
        self.x = Variable(const.shape)
        self.y = self.x
        self.mask = np.asarray(mask).astype(bool)
        self.const = np.asarray(const)
        scale = np.ones_like(self.const, dtype=np.float64)
        scale[self.mask, :] = 0.0
        self._jacobian = ScaleOp(self.x.size, scale.ravel())
        self._const = DeviceConstants()

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        y = _out(x, y)
        if x.numel() != self.const.size or y.numel() != self.const.size:
            raise ValueError("ConstrainOp: input/output have %d/%d entries, the constant %d"
                             % (x.numel(), y.numel(), self.const.size))
        if self.mask.size != self.const.shape[0]:
            raise ValueError("ConstrainOp: mask has %d entries for %d rows" % (self.mask.size, self.const.shape[0]))
        M = self.const.shape[1]
        c = self._const.get("c", self.const.ravel(), x.dtype)
        m = self._const.get_u8("mask", self.mask.ravel())
        _hip.call("constrain", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), M, _hip.ptr(c), _hip.ptr(m),
                  int(add), _hip.stream())
        if jacobian:
            return self._jacobian


class ShiftScaleOp(Operator):
    """T(x) = a*(x + b*shift), each of a, b, shift a scalar or an array
    (reference: affine.py:85-146)."""

    def __init__(self, N, shift, a, b):
        Operator.__init__(self)
        self.shift = shift
        self.a = a
        self.b = b
        self.x = Variable(N)
        self.y = Variable(N)
        self._jacobian = ScaleOp(N, self.a)
        self._const = DeviceConstants()

    def _arg(self, name, v, dtype, n):
        if isinstance(v, np.ndarray) and v.size == 1 and n != 1:
            v = float(v.ravel()[0])
        if isinstance(v, (np.ndarray, torch.Tensor)) and (v.size if isinstance(v, np.ndarray) else v.numel()) != n:
            raise ValueError("ShiftScaleOp: %s has %d entries, the input %d"
                             % (name, v.size if isinstance(v, np.ndarray) else v.numel(), n))
        if isinstance(v, np.ndarray):
            key = "%s:%d" % (name, id(v))
            if getattr(self, "_id_" + name, None) != key:
                self._const.drop(name)
                setattr(self, "_id_" + name, key)
            return _hip.ptr(self._const.get(name, v.ravel(), dtype)), 0.0
        if isinstance(v, torch.Tensor):
            return _hip.ptr(v), 0.0
        return None, float(v)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        y = _out(x, y)
        if y.numel() != x.numel():
            raise ValueError("ShiftScaleOp: input has %d entries, output %d" % (x.numel(), y.numel()))
        n = x.numel()
        sa, ss = self._arg("shift", self.shift, x.dtype, n)
        aa, as_ = self._arg("a", self.a, x.dtype, n)
        ba, bs = self._arg("b", self.b, x.dtype, n)
        b_zero = bs == 0 and ba is None
        s_zero = ss == 0 and sa is None
        if b_zero or s_zero:
            ba, bs = None, 0.0
        _hip.call("shiftscale", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), sa, ss, aa, as_, ba, bs, int(add),
                  _hip.stream())
        if jacobian:
            self._jacobian.fact = self.a
            return self._jacobian


class SSDProxOp(Operator):
    """prox_{tau G} for G = SSD(f, vol, mask): x -> (x + tau vol f)/(1 + tau vol)
    on unmasked points, identity elsewhere.  The reference expresses it as a
    ShiftScaleOp whose a, b arrays are recomputed and re-uploaded on every
    step change (reference: functionals/quad.py:45-55); here a and b are
    computed in the kernel from tau, which may be a device scalar."""

    def __init__(self, ssd, tau):
        Operator.__init__(self)
        self.ssd = ssd
        self.x = ssd.x
        self.y = ssd.x
        self.tau = StepArg(tau)

    def jacobian(self):
        """diag(a), a = 1/(1 + tau vol) on the data points and 1 elsewhere: the
        Jacobian of the reference's ShiftScaleOp form of this prox
        (reference: quad.py:45-55, affine.py:94, :133-135)."""
        s = self.ssd
        N, C = s.f.shape
        t = self.tau.tau
        if isinstance(t, tuple):  # (device FP64 tensor, index)
            t = float(t[0][t[1]].item())
        elif isinstance(t, torch.Tensor):
            t = t.detach().cpu().numpy()
        t = np.asarray(t, dtype=np.float64)
        tv = (np.full((N, C), float(t.ravel()[0])) if t.size == 1 else t.reshape(N, C)) * s.vol[None, :]
        tv[~s.mask, :] = 0.0
        key = tv.tobytes()
        if getattr(self, "_jac_key", None) != key:
            self._jac = ScaleOp(N * C, (1.0 / (1.0 + tv)).ravel())
            self._jac_key = key
        return self._jac

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        jac = self.jacobian() if jacobian else None
        y = _out(x, y)
        s = self.ssd
        N, C = s.f.shape
        f, vol, mask = s._dev(x.dtype)
        st = self.tau.step(x.dtype)
        _hip.call("ssd_prox", x.dtype, _hip.ptr(x), _hip.ptr(y), N, C, _hip.ptr(f), _hip.ptr(vol),
                  _hip.ptr(mask), _hip.ctypes.byref(st), int(add), _hip.stream())
        return jac


class L1DataProxOp(Operator):
    """prox of the TV-L1 data term G(x) = sum lbd_k |x - f| (soft threshold
    about f) or of its conjugate (conj=True: clip(x - tau f, -lbd, lbd)).
    No counterpart in the reference."""

    def __init__(self, fct, tau, conj):
        Operator.__init__(self)
        self.fct = fct
        self.x = fct.x
        self.y = fct.x
        self.conj = conj
        self.tau = StepArg(tau)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        if jacobian:
            raise NotImplementedError("L1DataProxOp: no Jacobian (the TV-L1 data term has no reference "
                                      "counterpart, and its prox is not differentiable at f +- t lbd)")
        y = _out(x, y)
        N, C = self.fct.f.shape
        f, lbd = self.fct._dev(x.dtype)
        st = self.tau.step(x.dtype)
        _hip.call("l1data_prox", x.dtype, _hip.ptr(x), _hip.ptr(y), N, C, _hip.ptr(f), _hip.ptr(lbd),
                  int(self.conj), _hip.ctypes.byref(st), int(add), _hip.stream())
