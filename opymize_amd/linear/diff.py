"""Finite-difference operators on Cartesian grids
(reference: opymize/linear/diff.py:67-249, diff.cu).

Layouts: image ``x`` is (npoints, C); vector field ``y`` is (npoints, D, C),
flat in C order, exactly as the reference's Variables.
"""
import numpy as np

from opymize_amd import LinOp, Variable, _hip
from opymize_amd._devcache import DeviceConstants

_SCHEMES = {"forward": _hip.SCHEME_FORWARD, "centered": _hip.SCHEME_CENTERED}


def _check_pair(x, y, nx, ny, what):
    if y is None:
        raise ValueError("%s needs an output array y" % what)
    if x.dtype != y.dtype:
        raise TypeError("%s: dtype mismatch %s vs %s" % (what, x.dtype, y.dtype))
    if x.numel() != nx or y.numel() != ny:
        raise ValueError("%s: sizes %d/%d, expected %d/%d" % (what, x.numel(), y.numel(), nx, ny))


class _DiffBase(LinOp):
    def _grid(self):
        return _hip.make_grid(self.imagedims, self.nchannels, self.imageh)

    def _w(self):
        return self._const.get("weights", np.asarray(self.weights, dtype=np.float64), _hip.torch_dtype("double"))


class GradientOp(_DiffBase):
    """y_t^k (+)= w_k/h_t * D_t x^k with the forward or centered scheme
    (reference: diff.py:67-144).  The centered scheme (the reference's
    default) averages the 2^(D-1) staggered forward differences of a cell;
    bottom/right boundary points of y are zero."""

    def __init__(self, imagedims, nchannels, scheme="centered", imageh=None, weights=None, adjoint=None):
        LinOp.__init__(self)
        if scheme not in _SCHEMES:
            raise ValueError("unknown scheme %r" % (scheme,))
        ndims = len(imagedims)
        npoints = int(np.prod(imagedims))
        self.imagedims = tuple(int(d) for d in imagedims)
        self.nchannels = int(nchannels)
        self.x = Variable((npoints, nchannels))
        self.y = Variable((npoints, ndims, nchannels))
        self.scheme = scheme
        self.imageh = np.ones(ndims) if imageh is None else np.asarray(imageh, dtype=np.float64)
        self.weights = np.ones(nchannels) if weights is None else np.asarray(weights, dtype=np.float64)
        self._const = DeviceConstants()
        if adjoint is None:
            self.adjoint = DivergenceOp(imagedims, nchannels, imageh=self.imageh, weights=self.weights,
                                        scheme=scheme, adjoint=self)
        else:
            self.adjoint = adjoint

    def _call_gpu(self, x, y=None, add=False):
        _check_pair(x, y, self.x.size, self.y.size, "GradientOp")
        g = self._grid()
        _hip.call("grad", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ctypes.byref(g), _hip.ptr(self._w()),
                  _SCHEMES[self.scheme], int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        g = self._grid()
        _hip.call("diff_rowwise", y.dtype, _hip.ptr(y, numel=self.y.size), _hip.ctypes.byref(g),
                  _hip.ptr(self._w()), _SCHEMES[self.scheme], 0, float(p), int(add), _hip.stream())


class DivergenceOp(_DiffBase):
    """Negative divergence with Dirichlet boundary, the adjoint of GradientOp:
    x^k (+)= -w_k div(y^k) (reference: diff.py:146-183)."""

    def __init__(self, imagedims, nchannels, imageh=None, weights=None, adjoint=None, scheme="centered"):
        LinOp.__init__(self)
        ndims = len(imagedims)
        npoints = int(np.prod(imagedims))
        self.imagedims = tuple(int(d) for d in imagedims)
        self.nchannels = int(nchannels)
        self.x = Variable((npoints, ndims, nchannels))
        self.y = Variable((npoints, nchannels))
        self.imageh = np.ones(ndims) if imageh is None else np.asarray(imageh, dtype=np.float64)
        self.weights = np.ones(nchannels) if weights is None else np.asarray(weights, dtype=np.float64)
        self._const = DeviceConstants()
        if adjoint is None:
            self.adjoint = GradientOp(imagedims, nchannels, scheme=scheme, imageh=self.imageh,
                                      weights=self.weights, adjoint=self)
        else:
            self.adjoint = adjoint
        self.scheme = self.adjoint.scheme

    def _call_gpu(self, x, y=None, add=False):
        _check_pair(x, y, self.x.size, self.y.size, "DivergenceOp")
        g = self._grid()
        _hip.call("div", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ctypes.byref(g), _hip.ptr(self._w()),
                  _SCHEMES[self.scheme], int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        g = self._grid()
        _hip.call("diff_rowwise", y.dtype, _hip.ptr(y, numel=self.y.size), _hip.ctypes.byref(g),
                  _hip.ptr(self._w()), _SCHEMES[self.scheme], 1, float(p), int(add), _hip.stream())


class LaplacianOp(LinOp):
    """Laplacian with 'neumann' (self-adjoint), 'curvature' or 'second-order'
    boundary conditions and their '_adj' adjoints (reference: diff.py:185-249)."""
    supported_bc = ["curvature", "curvature_adj", "second-order", "second-order_adj", "neumann"]

    def __init__(self, imagedims, nchannels, imageh=None, boundary="neumann", adjoint=None):
        LinOp.__init__(self)
        npoints = int(np.prod(imagedims))
        ndims = len(imagedims)
        self.imagedims = tuple(int(d) for d in imagedims)
        self.imageh = np.ones(ndims) if imageh is None else np.asarray(imageh, dtype=np.float64)
        self.bc = boundary
        self.nchannels = int(nchannels)
        self.x = Variable((npoints, nchannels))
        self.y = Variable((npoints, nchannels))
        if self.bc == "neumann":
            self.adjoint = self
        elif self.bc in self.supported_bc:
            if adjoint is None:
                adj_bc = self.bc[:-4] if self.bc.endswith("_adj") else "%s_adj" % self.bc
                self.adjoint = LaplacianOp(imagedims, nchannels, imageh=imageh, boundary=adj_bc, adjoint=self)
            else:
                self.adjoint = adjoint
        else:
            raise Exception("LaplacianOp: boundary condition %r is not one of %s" % (self.bc, list(self.supported_bc)))

    def _code(self):
        return ord(self.bc[0].upper()), int(self.bc.endswith("_adj"))

    def _call_gpu(self, x, y=None, add=False):
        _check_pair(x, y, self.x.size, self.y.size, "LaplacianOp")
        g = _hip.make_grid(self.imagedims, self.nchannels, self.imageh)
        bc, adj = self._code()
        _hip.call("laplacian", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ctypes.byref(g), bc, adj, int(add),
                  _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        g = _hip.make_grid(self.imagedims, self.nchannels, self.imageh)
        bc, adj = self._code()
        _hip.call("laplacian_rowwise", y.dtype, _hip.ptr(y, numel=self.y.size), _hip.ctypes.byref(g), bc, adj,
                  float(p), int(add), _hip.stream())
