"""Linear operators (reference: opymize/linear/__init__.py)."""
import logging

import numpy as np
import torch

from opymize_amd import LinOp, Variable, _hip
from opymize_amd.linear.diff import GradientOp, DivergenceOp, LaplacianOp
from opymize_amd.linear.einsum import (MatrixMult, MatrixMultR, MatrixOp, DiagMatrixMultR, TangledMatrixMultR,
                                       MatrixMultRBatched)
from opymize_amd.linear.indexed import IndexedMult, IndexedMultAdj
from opymize_amd.linear.scale import ZeroOp, ScaleOp, IdentityOp, NihilOp
from opymize_amd.linear.sparse import CsrOp


def _dev_norm(v):
    """FP64 Euclidean norm of a device vector (one reduction, one sync)."""
    out = torch.zeros(3, dtype=torch.float64, device=v.device)
    prm = _hip.FnParams()
    prm.c = v.data_ptr()
    _hip.call("fn_eval", v.dtype, _hip.FN_DOT, _hip.ptr(v), v.numel(), 1, _hip.ctypes.byref(prm),
              _hip.ptr(_hip.Workspace.get()), _hip.ptr(out), 0, _hip.stream())
    return float(np.sqrt(max(out[0].item(), 0.0)))


def normest(linop, xy=None, tol=1.0e-4, maxits=500, dtype=torch.float64):
    """Estimate the spectral norm of `linop` by power iteration on A^T A
    (reference: opymize/linear/__init__.py:11-46).

    Runs on the GPU.  The random sign vector and the restart vector come from
    numpy's global generator like the reference, so a seeded call starts from
    the same vector.
    """
    _hip.require_device()
    m, n = linop.y.size, linop.x.size
    itn = 0
    if xy is None:
        x = torch.empty(n, dtype=dtype, device="cuda")
        y = torch.empty(m, dtype=dtype, device="cuda")
    else:
        x, y = xy
    ysign = np.ones(m)
    ysign[np.random.randn(m) < 0] = -1
    y.copy_(torch.as_tensor(ysign, dtype=dtype))
    linop.adjoint(y, x)
    _hip.call("abs", x.dtype, _hip.ptr(x), _hip.ptr(x), n, _hip.stream())
    e = _dev_norm(x)
    if e < 1e-13:
        return e, itn
    _hip.call("scale", x.dtype, _hip.ptr(x), _hip.ptr(x), n, None, 1.0 / e, 0, _hip.stream())
    e0 = 0
    while abs(e - e0) > tol * e:
        e0 = e
        linop(x, y)
        normy = _dev_norm(y)
        if normy < 1e-13:
            y.copy_(torch.as_tensor(np.random.rand(m), dtype=dtype))
            normy = _dev_norm(y)
        linop.adjoint(y, x)
        normx = _dev_norm(x)
        e = normx / normy
        _hip.call("scale", x.dtype, _hip.ptr(x), _hip.ptr(x), n, None, 1.0 / normx, 0, _hip.stream())
        itn += 1
        if itn > maxits:
            logging.warning("normest: no convergence within %d power iterations (estimate %g)", maxits, e)
            break
    return e, itn


class BlockOp(LinOp):
    """A = [[A11, A12, ...], [A21, ...], ...]; y_i = sum_j A_ij x_j
    (reference: opymize/linear/__init__.py:48-111).  Blocks run as separate
    kernels accumulating into views of the flat output."""

    def __init__(self, blocks, adjoint=None):
        LinOp.__init__(self)
        x_sizes = [None] * len(blocks[0])
        y_sizes = [None] * len(blocks)
        for i, Bi in enumerate(blocks):
            for j, Bij in enumerate(Bi):
                try:
                    x_sizes[j] = (Bij.x.size,)
                    y_sizes[i] = (Bij.y.size,)
                except AttributeError:
                    pass
        self.x = Variable(*x_sizes)
        self.y = Variable(*y_sizes)
        for i, Bi in enumerate(blocks):
            for j, Bij in enumerate(Bi):
                if isinstance(Bij, (int, float)) and Bij == 0:
                    blocks[i][j] = ZeroOp(y_sizes[i][0], x_sizes[j][0])
        self.blocks = blocks
        for i, Bi in enumerate(self.blocks):
            for j, Bij in enumerate(Bi):
                # a mismatch would make the block kernels read or write past a view
                if Bij.y.size != self.blocks[i][0].y.size:
                    raise ValueError("BlockOp: block (%d,%d) has output size %d, row %d expects %d"
                                     % (i, j, Bij.y.size, i, self.blocks[i][0].y.size))
                if Bij.x.size != self.blocks[0][j].x.size:
                    raise ValueError("BlockOp: block (%d,%d) has input size %d, column %d expects %d"
                                     % (i, j, Bij.x.size, j, self.blocks[0][j].x.size))
        if adjoint is None:
            adj_bl = [[b[j].adjoint for b in self.blocks] for j in range(len(blocks[0]))]
            self.adjoint = BlockOp(adj_bl, adjoint=self)
        else:
            self.adjoint = adjoint

    def prepare_gpu(self, type_t="double"):
        for Bi in self.blocks:
            for Bij in Bi:
                Bij.prepare_gpu(type_t=type_t)

    def _call_gpu(self, x, y=None, add=False):
        yy = torch.empty_like(x) if y is None else y
        if not add:
            _hip.call("fill", yy.dtype, _hip.ptr(yy), yy.numel(), 0.0, _hip.stream())
        xs = self.x.vars(x)
        for j, yj in enumerate(self.y.vars(yy)):
            for i, xi in enumerate(xs):
                if isinstance(self.blocks[j][i], ZeroOp):
                    continue
                self.blocks[j][i](xi.reshape(-1), yj.reshape(-1), add=True)
        if y is None:
            _hip.call("copy", x.dtype, _hip.ptr(yy), _hip.ptr(x), x.numel(), _hip.stream())

    def _rowwise_gpu(self, y, p=1, add=False):
        if not add:
            _hip.call("fill", y.dtype, _hip.ptr(y), y.numel(), 0.0, _hip.stream())
        for j, yj in enumerate(self.y.vars(y)):
            for i in range(len(self.x.vars())):
                self.blocks[j][i].rowwise_lp(yj.reshape(-1), p=p, add=True)


__all__ = ["normest", "BlockOp", "GradientOp", "DivergenceOp", "LaplacianOp", "MatrixMult", "MatrixMultR",
           "MatrixOp", "DiagMatrixMultR", "TangledMatrixMultR", "MatrixMultRBatched", "IndexedMult", "IndexedMultAdj",
           "ZeroOp", "ScaleOp", "IdentityOp", "NihilOp", "CsrOp"]
