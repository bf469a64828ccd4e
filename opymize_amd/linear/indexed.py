"""Indexed products of lifting models (reference: opymize/linear/indexed.py:34-94).

IndexedMult gathers, per point i and region j, the L entries x[i, P[j, l]]
of the point's K values and applies the M x L matrix B[j]; IndexedMultAdj is
its adjoint (a scatter).  Both subtract, as the reference's kernels do
(indexed.cu:1-48): y = -B P x, accumulated into y with add=True."""
import numpy as np

from opymize_amd import LinOp, Variable, _hip
from opymize_amd._devcache import DeviceConstants
from opymize_amd.linear.einsum import _chk


def _merged(B, P, K):
    """Coefficient of x[., k] in output (j, ., m): -sum of B[j, m, l] over the l with P[j, l] = k; (J, M, K)."""
    J, M, L = B.shape
    C = np.zeros((J, M, K))
    for j in range(J):
        for l in range(L):
            C[j, :, P[j, l]] -= B[j, :, l]
    return C


class _IndexedBase(LinOp):
    def _setup(self, K, N, B, P):
        B = np.asarray(B, dtype=np.float64)
        P = np.asarray(P)
        if B.ndim != 3 or P.ndim != 2 or P.shape[0] != B.shape[0] or P.shape[1] != B.shape[2]:
            raise ValueError("IndexedMult: B must be (J, M, L) and P (J, L); got %s and %s" % (B.shape, P.shape))
        if P.size and (P.min() < 0 or P.max() >= K):
            raise ValueError("IndexedMult: indices of P must lie in [0, %d)" % K)
        self.K, self.N = int(K), int(N)
        self.B, self.P = B, P.astype(np.int32)
        self._const = DeviceConstants()

    def _launch(self, x, y, add, adj):
        J, M, L = self.B.shape
        Bd = self._const.get("B", self.B, x.dtype)
        Pd = self._const.get_i32("P", self.P)
        _hip.call("indexedmult", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(Bd), _hip.ptr(Pd), self.N, J, M, L,
                  self.K, int(adj), int(add), _hip.stream())


class IndexedMult(_IndexedBase):
    """(Ax)[j, i, m] -= sum_l B[j, m, l] x[i, P[j, l]] with x (N, K) and
    y (J, N, M) (reference: indexed.py:63-94)."""

    def __init__(self, K, N, B, P, adjoint=None):
        LinOp.__init__(self)
        self._setup(K, N, B, P)
        J, M, L = self.B.shape
        self.x = Variable((N, K))
        self.y = Variable((J, N, M))
        self.adjoint = IndexedMultAdj(K, N, P, B, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "IndexedMult")
        self._launch(x, y, add, adj=False)

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (j, i, m) has the entries _merged(B, P)[j, m, :] for every i
        rw = (np.abs(_merged(self.B, self.P, self.K)) ** p).sum(axis=2)  # (J, M)
        J, M = rw.shape
        r = self._const.get("rowwise%g" % p, rw.ravel(), y.dtype)
        for j in range(J):
            _hip.call("tile", y.dtype, _hip.ptr(y.view(-1)[j * self.N * M:(j + 1) * self.N * M]), self.N * M,
                      _hip.ptr(r[j * M:(j + 1) * M]), M, 1, int(add), _hip.stream())


class IndexedMultAdj(_IndexedBase):
    """(Ax)[i, P[j, l]] -= sum_m B[j, m, l] x[j, i, m] with x (J, N, M) and
    y (N, K): the adjoint of IndexedMult (reference: indexed.py:34-61)."""

    def __init__(self, K, N, P, B, adjoint=None):
        LinOp.__init__(self)
        self._setup(K, N, B, P)
        J, M, L = self.B.shape
        self.x = Variable((J, N, M))
        self.y = Variable((N, K))
        self.adjoint = IndexedMult(K, N, B, P, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "IndexedMultAdj")
        self._launch(x, y, add, adj=True)

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (i, k) has the entries _merged(B, P)[:, :, k] for every i
        rk = (np.abs(_merged(self.B, self.P, self.K)) ** p).sum(axis=(0, 1))  # (K,)
        r = self._const.get("rowwise%g" % p, rk, y.dtype)
        _hip.call("tile", y.dtype, _hip.ptr(y, numel=self.y.size), self.y.size, _hip.ptr(r), r.numel(), 1, int(add),
                  _hip.stream())
