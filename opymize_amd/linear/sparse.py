"""GPU path for linear operators given only as a matrix.

The reference evaluates such LinOps with ``spmat.dot`` on the CPU and has no
GPU path for them (reference: opymize/__init__.py:92-108).  Here the matrix
is converted once to CSR (int64 indices) and applied with a HIP kernel; the
adjoint uses the transposed CSR.
"""
import numpy as np
import scipy.sparse as sp
import torch

from opymize_amd import LinOp, Variable, _hip
from opymize_amd._devcache import DeviceConstants


class CsrOp(LinOp):
    """y (+)= M x for a scipy sparse (or dense) matrix M."""

    def __init__(self, mat, adjoint=None):
        LinOp.__init__(self)
        m = sp.csr_matrix(mat, dtype=np.float64)
        m.sum_duplicates()
        m.sort_indices()
        self.spmat = m
        self.y = Variable(m.shape[0])
        self.x = Variable(m.shape[1])
        self._const = DeviceConstants()
        self.adjoint = CsrOp(m.T, adjoint=self) if adjoint is None else adjoint

    def _dev(self, dtype):
        m = self.spmat
        indptr = self._const.get("indptr", m.indptr.astype(np.int64), torch.int64)
        indices = self._const.get("indices", m.indices.astype(np.int64), torch.int64)
        data = self._const.get("data", m.data, dtype)
        return indptr, indices, data

    def _call_gpu(self, x, y=None, add=False):
        if y is None:
            raise ValueError("CsrOp needs an output array y")
        if x.numel() != self.x.size or y.numel() != self.y.size or x.dtype != y.dtype:
            raise ValueError("CsrOp: size/dtype mismatch")
        ip, ix, d = self._dev(x.dtype)
        _hip.call("csr_spmv", x.dtype, self.y.size, _hip.ptr(ip), _hip.ptr(ix), _hip.ptr(d), _hip.ptr(x),
                  _hip.ptr(y), int(add), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        ip, _, d = self._dev(y.dtype)
        _hip.call("csr_rowwise", y.dtype, self.y.size, _hip.ptr(ip), _hip.ptr(d), float(p), _hip.ptr(y), int(add),
                  _hip.stream())
