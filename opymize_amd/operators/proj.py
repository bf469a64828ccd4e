"""Projections (reference: opymize/operators/proj.py)."""
import numpy as np
import torch

from opymize_amd import LinOp, Operator, Variable, _hip
from opymize_amd._devcache import DeviceConstants


def _inplace_target(x, y):
    """In-place kernels act on y after copying x into it."""
    if y is None or y.data_ptr() == x.data_ptr():
        return x
    _hip.call("copy", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), _hip.stream())
    return y


class _PointJacobian(LinOp):
    """Jacobian of a pointwise projection at the point it was last applied
    to: keeps a device copy of that input (the projections run in place)."""

    def __init__(self, N):
        LinOp.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self.adjoint = self  # projections have symmetric Jacobians
        self.xbar = None

    def update(self, xbar):
        if self.xbar is None or self.xbar.dtype != xbar.dtype or self.xbar.numel() != xbar.numel() \
                or self.xbar.device != xbar.device:
            self.xbar = torch.empty(xbar.numel(), dtype=xbar.dtype, device=xbar.device)
        _hip.call("copy", xbar.dtype, _hip.ptr(xbar), _hip.ptr(self.xbar), xbar.numel(), _hip.stream())
        return self

    def _point(self, x, y):
        """The stored point, after checking x and the output y against it."""
        if self.xbar is None:
            raise RuntimeError("%s: no point to linearize at (apply the projection with jacobian=True)"
                               % type(self).__name__)
        for t in (x, y):
            if t.numel() != self.xbar.numel() or t.dtype != self.xbar.dtype:
                raise ValueError("%s: %d entries of %s for a Jacobian at %d entries of %s"
                                 % (type(self).__name__, t.numel(), t.dtype, self.xbar.numel(), self.xbar.dtype))
        return self.xbar


class SignProjJacobian(_PointJacobian):
    """Jacobian of PosProj / NegProj at xbar: keeps x where xbar > 0 (< 0),
    zero elsewhere (reference: proj.py:25, :38-40: NihilOp with keep taken
    from the input)."""

    def __init__(self, N, positive):
        _PointJacobian.__init__(self, N)
        self.positive = bool(positive)

    def _call_gpu(self, x, y=None, add=False):
        y = x if y is None else y
        xb = self._point(x, y)
        _hip.call("signproj_jac", x.dtype, _hip.ptr(xb), _hip.ptr(x), _hip.ptr(y), x.numel(), int(self.positive),
                  int(add), _hip.stream())


class PosProj(Operator):
    """T(x) = max(0, x) (reference: proj.py:19-52)."""

    def __init__(self, N):
        Operator.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self._jacobian = SignProjJacobian(N, True)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        jac = self._jacobian.update(x) if jacobian else None
        y = x if y is None else y
        _hip.call("signproj", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), 1, int(add), _hip.stream())
        return jac


class NegProj(Operator):
    """T(x) = min(0, x) (reference: proj.py:54-85)."""

    def __init__(self, N):
        Operator.__init__(self)
        self.x = Variable(N)
        self.y = Variable(N)
        self._jacobian = SignProjJacobian(N, False)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        jac = self._jacobian.update(x) if jacobian else None
        y = x if y is None else y
        _hip.call("signproj", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), 0, int(add), _hip.stream())
        return jac


class L12ProjJacobian(_PointJacobian):
    """Jacobian of the Frobenius L1NormsProj at xbar (reference:
    proj.py:228-275): identity on points inside the ball, and
    (lbd/|xbar_i|)(I - u u^T), u = xbar_i/|xbar_i|, outside."""

    def __init__(self, N, M, lbd):
        _PointJacobian.__init__(self, N * M[0] * M[1])
        self.x = Variable((N, M[0] * M[1]))
        self.y = self.x
        self.N, self.M, self.lbd = int(N), tuple(M), float(lbd)

    def _call_gpu(self, x, y=None, add=False):
        y = x if y is None else y
        xb = self._point(x, y)
        _hip.call("l12proj_jac", x.dtype, _hip.ptr(xb), _hip.ptr(x), _hip.ptr(y), self.N, self.M[0] * self.M[1],
                  self.lbd, int(add), _hip.stream())


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
        # a ball of vectors is the same for both norms
        frob = matrixnorm == "frobenius" or min(M) == 1
        self._jacobian = L12ProjJacobian(self.N, M, self.lbd) if frob else None

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        jac = None
        if jacobian:
            # the reference has the Jacobian of the Frobenius ball only (proj.py:211)
            if self._jacobian is None:
                raise NotImplementedError("L1NormsProj: no Jacobian for the spectral ball")
            jac = self._jacobian.update(x)
        t = _inplace_target(x, y)
        _hip.call("l1norms_proj", t.dtype, _hip.ptr(t), self.N, self.M[0], self.M[1], self.lbd,
                  int(self.matrixnorm == "spectral"), _hip.stream())
        return jac


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


def epigraph_Ab(I, J, v, b):
    """Constraint system of the epigraph projection as a block-diagonal CSR
    matrix A and right-hand side b, one block per (region j, function i) in
    C order, rows (v[l], -1) and b[i, l] for l = J[j, k] with I[i, l]
    (reference: proj.py:376-386)."""
    import scipy.sparse as sp
    I = np.asarray(I, dtype=bool)
    J = np.asarray(J)
    v = np.asarray(v, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nfuns = I.shape[0]
    nregions, nsub = J.shape
    ndim = v.shape[1]
    rows, bs = [], []
    for j in range(nregions):
        for i in range(nfuns):
            ls = [l for l in J[j] if I[i, l]]
            blk = -np.ones((len(ls), ndim + 1))
            blk[:, :-1] = v[ls]
            rows.append(blk)
            bs.append(b[i, ls])
    A = sp.block_diag(rows, format="csr") if rows else sp.csr_matrix((0, 0))
    return A, np.concatenate(bs) if bs else np.zeros(0)


class EpigraphProj(Operator):
    """T(x)[j,i] = projection of x[j,i] onto the epigraph of f[i]_j*, the
    conjugate of the convex restriction of the piecewise linear f[i] (values
    b[i,k] at the points v[k], k in I[i]) to the region J[j]:

        minimize 0.5*|y - x[j,i]|^2
        s.t. <v[l], y[:-1]> - y[-1] <= b[i,l] for l in J[j] with I[i,l]

    (reference: proj.py:388-470, proj.cu:222-831).  Solved per (j, i) on the
    GPU by an active-set method with at most ndim+1 active constraints and
    ndim*nsubpoints iterations, ndim <= 3."""

    def __init__(self, I, J, v, b, Ab=None):
        Operator.__init__(self)
        I = np.asarray(I, dtype=bool)
        J = np.asarray(J)
        v = np.asarray(v, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        assert I.shape == b.shape
        assert I.shape[1] == v.shape[0]
        nfuns, npoints = I.shape
        nregions, nsubpoints = J.shape
        ndim = v.shape[1]
        if not 1 <= ndim <= 3:
            raise ValueError("EpigraphProj supports 1 <= ndim <= 3, got %d" % ndim)
        if J.size and (J.min() < 0 or J.max() >= npoints):
            raise ValueError("J indexes points outside [0, %d)" % npoints)
        self.I, self.J, self.v, self.b = I, J, v, b
        self._Ab = Ab
        self.x = Variable((nregions, nfuns, ndim + 1))
        self.y = self.x
        self.maxiter = ndim * nsubpoints
        self.tol = 1e-9
        self._const = DeviceConstants()

    @property
    def Ab(self):
        if self._Ab is None:
            self._Ab = epigraph_Ab(self.I, self.J, self.v, self.b)
        return self._Ab

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        assert not jacobian
        t = _inplace_target(x, y)
        nfuns, npoints = self.I.shape
        nregions, nsub = self.J.shape
        dI = self._const.get_u8("I", self.I)
        dJ = self._const.get("J", self.J.astype(np.int32), torch.int32)
        dv = self._const.get("v", self.v, t.dtype)
        db = self._const.get("b", self.b, t.dtype)
        _hip.call("epigraph_proj", t.dtype, _hip.ptr(t), _hip.ptr(dI), _hip.ptr(dJ), _hip.ptr(dv), _hip.ptr(db),
                  self.v.shape[1], nfuns, nregions, npoints, nsub, self.maxiter, self.tol, _hip.stream())


class ClipOp(Operator):
    """T(x) = min(max(x, lo), hi): projection onto the box [lo, hi]^N (no
    reference counterpart; the prox of BoxFct)."""

    def __init__(self, N, lo=0.0, hi=1.0):
        Operator.__init__(self)
        if not lo <= hi:
            raise ValueError("ClipOp needs lo <= hi, got %g > %g" % (lo, hi))
        self.x = Variable(N)
        self.y = self.x
        self.lo, self.hi = float(lo), float(hi)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        assert not jacobian
        y = x if y is None else y
        if x.numel() != self.x.size or y.numel() != self.x.size:
            raise ValueError("ClipOp: %d -> %d entries, expected %d" % (x.numel(), y.numel(), self.x.size))
        _hip.call("clip", x.dtype, _hip.ptr(x), _hip.ptr(y), x.numel(), self.lo, self.hi, _hip.stream())


class SimplexProj(Operator):
    """T(x)[i] = projection of the i-th point's K values onto the probability
    simplex {v >= 0, sum v = 1} (no reference counterpart; the prox of
    SimplexFct).  Michelot's finite algorithm, one thread per point."""

    def __init__(self, N, K):
        Operator.__init__(self)
        if not 1 <= K <= 32:
            raise ValueError("SimplexProj supports 1 <= K <= 32, got %d" % K)
        self.N, self.K = int(N), int(K)
        self.x = Variable((self.N, self.K))
        self.y = self.x

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        assert not add
        assert not jacobian
        if x.numel() != self.x.size:
            raise ValueError("SimplexProj: %d entries, expected %d" % (x.numel(), self.x.size))
        t = _inplace_target(x, y)
        _hip.call("simplex_proj", t.dtype, _hip.ptr(t), self.N, self.K, _hip.stream())
