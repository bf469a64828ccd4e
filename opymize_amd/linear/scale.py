"""Trivial linear operators (reference: opymize/linear/scale.py:14-115)."""
import numpy as np

from opymize_amd import LinOp, Variable, _hip
from opymize_amd._devcache import DeviceConstants


def _fill(y, v):
    _hip.call("fill", y.dtype, _hip.ptr(y), y.numel(), float(v), _hip.stream())


class ZeroOp(LinOp):
    """(Ax)_i = 0 (reference: scale.py:14-31)."""

    def __init__(self, M, N=None, adjoint=None):
        LinOp.__init__(self)
        N = M if N is None else N
        self.x = Variable(N)
        self.y = Variable(M)
        self.adjoint = ZeroOp(N, M, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        if not add:
            _fill(x if y is None else y, 0.0)

    def _rowwise_gpu(self, y, p=1, add=False):
        if not add:
            _fill(y, 0.0)


class ScaleOp(LinOp):
    """(Ax)_i = fact*x_i, fact scalar or array (reference: scale.py:33-74)."""

    def __init__(self, N, fact):
        LinOp.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self.adjoint = self
        self.fact = fact
        self._const = DeviceConstants()

    def _fact_args(self, dtype):
        if isinstance(self.fact, np.ndarray):
            if getattr(self, "_fact_id", None) != id(self.fact):
                self._const.drop("fact")  # fact was replaced since the last upload
                self._fact_id = id(self.fact)
            return _hip.ptr(self._const.get("fact", self.fact.ravel(), dtype)), 0.0
        return None, float(self.fact)

    def _call_gpu(self, x, y=None, add=False):
        y = x if y is None else y
        fa, fs = self._fact_args(x.dtype)
        _hip.call("scale", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), fa, fs, int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        if isinstance(self.fact, np.ndarray):
            r = self._const.get("rowwise%g" % p, np.abs(self.fact.ravel()) ** p, y.dtype)
            _hip.call("tile", y.dtype, _hip.ptr(y), y.numel(), _hip.ptr(r), r.numel(), 1, int(add), _hip.stream())
        else:
            v = abs(float(self.fact)) ** p
            if add:
                vv = self._const.get("rowwise%g" % p, np.array([v]), y.dtype)
                _hip.call("tile", y.dtype, _hip.ptr(y), y.numel(), _hip.ptr(vv), 1, 1, 1, _hip.stream())
            else:
                _fill(y, v)


class IdentityOp(LinOp):
    """(Ax)_i = x_i (reference: scale.py:76-94)."""

    def __init__(self, N):
        LinOp.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self.adjoint = self

    def _call_gpu(self, x, y=None, add=False):
        if y is None and not add:
            return
        y = x if y is None else y
        _hip.call("scale", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), None, 1.0, int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        if add:
            one = DeviceConstants().get("one", np.ones(1), y.dtype)
            _hip.call("tile", y.dtype, _hip.ptr(y), y.numel(), _hip.ptr(one), 1, 1, 1, _hip.stream())
        else:
            _fill(y, 1.0)


class NihilOp(LinOp):
    """(Ax)_i = x_i if keep_i else 0 (reference: scale.py:96-115; the
    reference has no GPU path for it)."""

    def __init__(self, N, keep):
        LinOp.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self.adjoint = self
        self.keep = np.asarray(keep).astype(bool)
        self._const = DeviceConstants()

    def _call_gpu(self, x, y=None, add=False):
        y = x if y is None else y
        k = self._const.get_u8("keep", self.keep.ravel())
        _hip.call("nihil", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), _hip.ptr(k), int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        r = self._const.get("rowwise", self.keep.ravel().astype(np.float64), y.dtype)
        _hip.call("tile", y.dtype, _hip.ptr(y), y.numel(), _hip.ptr(r), r.numel(), 1, int(add), _hip.stream())
