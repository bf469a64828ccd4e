"""Dense per-point matrix products (reference: opymize/linear/einsum.py:17-128)."""
import numpy as np

from opymize_amd import LinOp, Variable, _hip
from opymize_amd._devcache import DeviceConstants


def _chk(x, y, nx, ny, what):
    if y is None:
        raise ValueError("%s needs an output array y" % what)
    if x.dtype != y.dtype or x.numel() != nx or y.numel() != ny:
        raise ValueError("%s: got %s[%d] -> %s[%d], expected %d -> %d"
                         % (what, x.dtype, x.numel(), y.dtype, y.numel(), nx, ny))


class MatrixMult(LinOp):
    """Label-major product y_ji = sum_k A_jk x_ki (trans: A_kj)
    (reference: einsum.py:17-56).

    The reference sizes x as (A.shape[0], N) and y as (A.shape[1], N), which
    only agrees with its own formula for square A; here x is (K, N) and y is
    (J, N) for the (J, K) matrix of the formula, identical for square A."""

    def __init__(self, N, A, trans=False, adjoint=None):
        LinOp.__init__(self)
        A = np.asarray(A, dtype=np.float64)
        k, j = (0, 1) if trans else (1, 0)
        self.x = Variable((A.shape[k], N))
        self.y = Variable((A.shape[j], N))
        self.N = int(N)
        self.trans = trans
        self.A = A
        self._const = DeviceConstants()
        self.adjoint = MatrixMult(N, A, trans=not trans, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "MatrixMult")
        K = self.x[0]["shape"][0]
        J = self.y[0]["shape"][0]
        Ad = self._const.get("A", self.A, x.dtype)
        _hip.call("matmul", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(Ad), self.N, J, K, int(self.trans),
                  int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (j, i) has the entries A(j, :) for every i
        rows = np.abs(self.A.T if self.trans else self.A) ** p
        r = self._const.get("rowwise%g" % p, rows.sum(axis=1), y.dtype)
        _hip.call("tile", y.dtype, _hip.ptr(y, numel=self.y.size), self.y.size, _hip.ptr(r), r.numel(), self.N,
                  int(add), _hip.stream())


class MatrixMultR(LinOp):
    """Point-major product y_ij = sum_k x_ik A_kj (trans: A_jk)
    (reference: einsum.py:58-97)."""

    def __init__(self, N, A, trans=False, adjoint=None):
        LinOp.__init__(self)
        A = np.asarray(A, dtype=np.float64)
        k, j = (1, 0) if trans else (0, 1)
        self.x = Variable((N, A.shape[k]))
        self.y = Variable((N, A.shape[j]))
        self.N = int(N)
        self.trans = trans
        self.A = A
        self._const = DeviceConstants()
        self.adjoint = MatrixMultR(N, A, trans=not trans, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "MatrixMultR")
        K = self.x[0]["shape"][1]
        J = self.y[0]["shape"][1]
        Ad = self._const.get("A", self.A, x.dtype)
        _hip.call("matmul_r", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(Ad), self.N, J, K, int(self.trans),
                  int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (i, j) has the entries A(:, j) for every i
        cols = np.abs(self.A.T if self.trans else self.A) ** p
        r = self._const.get("rowwise%g" % p, cols.sum(axis=0), y.dtype)
        _hip.call("tile", y.dtype, _hip.ptr(y, numel=self.y.size), self.y.size, _hip.ptr(r), r.numel(), 1,
                  int(add), _hip.stream())


class MatrixOp(MatrixMult):
    """y_i = sum_j A_ij x_j (reference: einsum.py:99-102)."""

    def __init__(self, A):
        MatrixMult.__init__(self, 1, A)


class DiagMatrixMultR(LinOp):
    """y_ik = x_ik A_k (reference: einsum.py:104-128)."""

    def __init__(self, N, A):
        LinOp.__init__(self)
        A = np.asarray(A, dtype=np.float64).ravel()
        self.x = Variable((N, A.size))
        self.y = Variable((N, A.size))
        self.N = int(N)
        self.A = A
        self.adjoint = self
        self._const = DeviceConstants()

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "DiagMatrixMultR")
        Ad = self._const.get("A", self.A, x.dtype)
        _hip.call("diag_matmul_r", x.dtype, _hip.ptr(x), _hip.ptr(y), self.N, self.A.size, _hip.ptr(Ad), int(add),
                  _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        r = self._const.get("rowwise%g" % p, np.abs(self.A) ** p, y.dtype)
        _hip.call("tile", y.dtype, _hip.ptr(y, numel=self.y.size), self.y.size, _hip.ptr(r), r.numel(), 1,
                  int(add), _hip.stream())


class TangledMatrixMultR(LinOp):
    """Entangled product from the right: y_mik = sum_jl x_jil A_jlmk, i.e.
    y_(i)(mk) = sum_(jl) x_(i)(jl) A_(jl)(mk); trans: A_mkjl, the adjoint
    (reference: einsum.py:130-173)."""

    def __init__(self, N, A, trans=False, adjoint=None):
        LinOp.__init__(self)
        A = np.asarray(A, dtype=np.float64)
        if A.ndim != 4:
            raise ValueError("TangledMatrixMultR: A must have 4 axes, got %s" % (A.shape,))
        j, m = (2, 0) if trans else (0, 2)
        self.x = Variable((A.shape[j], N, A.shape[j + 1]))
        self.y = Variable((A.shape[m], N, A.shape[m + 1]))
        self.N = int(N)
        self.trans = trans
        self.A = A
        self._const = DeviceConstants()
        self.adjoint = TangledMatrixMultR(N, A, trans=not trans, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "TangledMatrixMultR")
        J, _, L = self.x[0]["shape"]
        M, _, K = self.y[0]["shape"]
        Ad = self._const.get("A", self.A, x.dtype)
        _hip.call("tangled_matmul_r", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(Ad), self.N, J, L, M, K,
                  int(self.trans), int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (m, i, k) has the entries A(:, :, m, k) for every i
        Am = np.abs(self.A) ** p
        rw = Am.sum(axis=(2, 3)) if self.trans else Am.sum(axis=(0, 1))  # (M, K)
        M, K = rw.shape
        r = self._const.get("rowwise%g" % p, rw.ravel(), y.dtype)
        for m in range(M):
            _hip.call("tile", y.dtype, _hip.ptr(y.view(-1)[m * self.N * K:(m + 1) * self.N * K]), self.N * K,
                      _hip.ptr(r[m * K:(m + 1) * K]), K, 1, int(add), _hip.stream())


class MatrixMultRBatched(LinOp):
    """Batched product from the right: y_jlk = sum_m x_jlm A_jmk with the
    point index l; trans: A_jkm, the adjoint (reference: einsum.py:175-214)."""

    def __init__(self, N, A, trans=False, adjoint=None):
        LinOp.__init__(self)
        A = np.asarray(A, dtype=np.float64)
        if A.ndim != 3:
            raise ValueError("MatrixMultRBatched: A must have 3 axes, got %s" % (A.shape,))
        m, k = (2, 1) if trans else (1, 2)
        self.x = Variable((A.shape[0], N, A.shape[m]))
        self.y = Variable((A.shape[0], N, A.shape[k]))
        self.N = int(N)
        self.trans = trans
        self.A = A
        self._const = DeviceConstants()
        self.adjoint = MatrixMultRBatched(N, A, trans=not trans, adjoint=self) if adjoint is None else adjoint

    def _call_gpu(self, x, y=None, add=False):
        _chk(x, y, self.x.size, self.y.size, "MatrixMultRBatched")
        J, _, M = self.x[0]["shape"]
        K = self.y[0]["shape"][2]
        Ad = self._const.get("A", self.A, x.dtype)
        _hip.call("matmul_r_batched", x.dtype, _hip.ptr(x), _hip.ptr(y), _hip.ptr(Ad), self.N, J, M, K,
                  int(self.trans), int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        # row (j, i, k) has the entries A(j, :, k) for every i
        Am = np.abs(self.A) ** p
        rw = Am.sum(axis=2) if self.trans else Am.sum(axis=1)  # (J, K)
        J, K = rw.shape
        r = self._const.get("rowwise%g" % p, rw.ravel(), y.dtype)
        for j in range(J):
            _hip.call("tile", y.dtype, _hip.ptr(y.view(-1)[j * self.N * K:(j + 1) * self.N * K]), self.N * K,
                      _hip.ptr(r[j * K:(j + 1) * K]), K, 1, int(add), _hip.stream())
