"""Projections (reference: opymize/operators/proj.py)."""
import numpy as np
import torch

from opymize_amd import Operator, Variable, _hip
from opymize_amd._devcache import DeviceConstants


def _inplace_target(x, y):
    """In-place kernels act on y after copying x into it."""
    if y is None or y.data_ptr() == x.data_ptr():
        return x
    _hip.call("copy", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), _hip.stream())
    return y


class PosProj(Operator):
    """T(x) = max(0, x) (reference: proj.py:19-52)."""

    def __init__(self, N):
        Operator.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        if jacobian:
            raise NotImplementedError("jacobians of projections are not part of the PDHG hot path")
        y = x if y is None else y
        _hip.call("signproj", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), 1, int(add), _hip.stream())


class NegProj(Operator):
    """T(x) = min(0, x) (reference: proj.py:54-85)."""

    def __init__(self, N):
        Operator.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        if jacobian:
            raise NotImplementedError("jacobians of projections are not part of the PDHG hot path")
        y = x if y is None else y
        _hip.call("signproj", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), 0, int(add), _hip.stream())


class L1NormsProj(Operator):
    """Per-point projection onto the ball {|x_i| <= lbd} for the Frobenius or
    spectral norm of the (M1 x M2) block x_i (reference: proj.py:170-226,
    proj.cu:3-107).

    Blocks are read column by column (entry (r, c) at offset r + c*M1), the
    indexing of the reference's GPU kernel: with M = (C, D) each block is the
    Jacobian of the gradient's (D, C) per-point layout.  The Frobenius ball
    does not depend on that choice."""

    def __init__(self, N, M, lbd, matrixnorm):
        Operator.__init__(self)
        assert matrixnorm in ["frobenius", "spectral"]
        M = tuple(int(m) for m in M)
        if matrixnorm == "spectral" and min(M) > 2:
            raise ValueError("spectral projection needs a block side of length <= 2, got %s" % (M,))
        self.M = M
        self.N = int(N)
        self.x = Variable((self.N,) + M)
        self.y = self.x
        self.lbd = float(lbd)
        self.matrixnorm = matrixnorm

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        assert not jacobian
        t = _inplace_target(x, y)
        _hip.call("l1norms_proj", t.dtype, _hip.ptr(t), self.N, self.M[0], self.M[1], self.lbd,
                  int(self.matrixnorm == "spectral"), _hip.stream())


class QuadEpiProj(Operator):
    """T(z)[i] = projection of z[i] = (x1, x2) onto the epigraph of
    f_i(x) = 0.5*alph*|x - shift1[i]|^2 + shift2[i], optionally truncated at
    |x - shift1[i]| <= lbd (reference: proj.py:278-374, proj.cu:110-219).
    With b, c given, shift1 = -b/alph and shift2 = c - 0.5/alph*|b|^2."""

    def __init__(self, N, M, alph=1.0, shift=None, b=None, c=None, lbd=np.inf):
        Operator.__init__(self)
        assert lbd > 0
        assert alph > 0
        if M > _hip.OPY_QUADEPI_MAXM:
            raise ValueError("QuadEpiProj supports M <= %d" % _hip.OPY_QUADEPI_MAXM)
        self.N, self.M = int(N), int(M)
        self.x = Variable((self.N, self.M + 1))
        self.y = self.x
        self.lbd = lbd
        self.alph = float(alph)
        self.shift = shift
        if not (b is None and c is None):
            assert self.shift is None
            self.b = np.zeros((N, M)) if b is None else np.asarray(b, dtype=np.float64)
            self.c = np.zeros((N,)) if c is None else np.asarray(c, dtype=np.float64)
            self.shift = np.zeros((N, M + 1))
            self.shift[:, :-1] = -self.b / self.alph
            self.shift[:, -1] = -(0.5 / self.alph * (self.b ** 2).sum(axis=-1) - self.c)
        self._const = DeviceConstants()

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        assert not jacobian
        t = _inplace_target(x, y)
        sh = None
        if self.shift is not None:
            sh = self._const.get("shift", np.asarray(self.shift).ravel(), x.dtype)
        _hip.call("quadepi_proj", t.dtype, _hip.ptr(t), self.N, self.M, self.alph, _hip.ptr(sh),
                  float(self.lbd), _hip.stream())
