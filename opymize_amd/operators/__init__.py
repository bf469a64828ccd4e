"""Nonlinear operators (prox maps) (reference: opymize/operators/__init__.py)."""
from opymize_amd import Operator, Variable
from opymize_amd.operators.affine import (ConstOp, ConstrainOp, ShiftScaleOp, SSDProxOp, L1DataProxOp,
                                          StepArg)
from opymize_amd.operators.proj import (PosProj, NegProj, L1NormsProj, QuadEpiProj, EpigraphProj, epigraph_Ab,
                                       ClipOp, SimplexProj)


class SplitOp(Operator):
    """T(x1, x2, ...) = (T1(x1), T2(x2), ...) (reference: operators/__init__.py:6-27)."""

    def __init__(self, ops):
        Operator.__init__(self)
        self.x = Variable(*[(T.x.size,) for T in ops])
        self.y = Variable(*[(T.y.size,) for T in ops])
        self.ops = ops

    def prepare_gpu(self, type_t="double"):
        for op in self.ops:
            op.prepare_gpu(type_t=type_t)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        X = self.x.vars(x)
        Y = [None] * len(X) if y is None else self.y.vars(y)
        results = [[0] * len(self.ops) for _ in self.ops]
        for i, (T, xi, yi) in enumerate(zip(self.ops, X, Y)):
            results[i][i] = T(xi.reshape(-1), None if yi is None else yi.reshape(-1), add=add, jacobian=jacobian)
        if jacobian:
            from opymize_amd.linear import BlockOp
            return BlockOp(results)


__all__ = ["ClipOp", "SimplexProj", "SplitOp", "ConstOp", "ConstrainOp", "ShiftScaleOp", "SSDProxOp", "L1DataProxOp", "StepArg",
           "PosProj", "NegProj", "L1NormsProj", "QuadEpiProj"]
