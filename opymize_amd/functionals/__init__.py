"""Functionals (reference: opymize/functionals/__init__.py)."""
import numpy as np
import torch

from opymize_amd import Functional, Variable
from opymize_amd.operators import SplitOp
from opymize_amd.functionals.affine import AffineFct, ConstFct, ZeroFct, MaskedAffineFct
from opymize_amd.functionals.indicators import (IndicatorFct, ConstrainFct, PositivityFct, NegativityFct, BoxFct,
                                                BoxSupp, SimplexFct, SimplexSupp)
from opymize_amd.functionals.l1norms import L1Norms, L1NormsConj
from opymize_amd.functionals.quad import SSD, L1Data, L1DataConj
from opymize_amd.functionals.epigraph import (QuadEpiInd, QuadEpiSupp, QuadSupport, HuberPerspective,
                                              TruncQuadEpiInd, quad_dual_coefficients, EpigraphInd,
                                              EpigraphFct, EpigraphSupp, EpigraphSupportFct)


class SplitSum(Functional):
    """F(x1, x2, ...) = F1(x1) + F2(x2) + ...; value = sum, infeas = max
    (reference: functionals/__init__.py:10-38)."""

    def __init__(self, fcts, conj=None):
        Functional.__init__(self)
        self.x = Variable(*[(F.x.size,) for F in fcts])
        self.fcts = fcts
        self.conj = SplitSum([F.conj for F in fcts], conj=self) if conj is None else conj

    def nslots(self):
        return sum(F.nslots() for F in self.fcts)

    def precond_blocks(self):
        out, off = [], 0
        for F in self.fcts:
            out += [(off + o, nb, w) for o, nb, w in F.precond_blocks()]
            off += F.x.size
        return out

    def _gpu_eval(self, x, out):
        o = 0
        for F, xi in zip(self.fcts, self.x.vars(x)):
            k = F.nslots()
            F._gpu_eval(xi.reshape(-1), out[3 * o:3 * (o + k)])
            o += k

    def _gpu_finish(self, v):
        vals, infs, o = [], [], 0
        for F in self.fcts:
            k = F.nslots()
            val, inf = F._gpu_finish(v[3 * o:3 * (o + k)])
            vals.append(val)
            infs.append(inf)
            o += k
        return sum(vals), max(infs)

    def prox(self, tau):
        per_entry = (isinstance(tau, np.ndarray) and tau.size > 1) or \
            (isinstance(tau, torch.Tensor) and tau.numel() > 1)
        if per_entry:
            taus = self.x.vars(tau)
            return SplitOp([F.prox(t.reshape(-1)) for F, t in zip(self.fcts, taus)])
        return SplitOp([F.prox(tau) for F in self.fcts])


__all__ = ["SplitSum", "AffineFct", "ConstFct", "ZeroFct", "MaskedAffineFct", "IndicatorFct", "ConstrainFct",
           "BoxFct", "BoxSupp", "SimplexFct", "SimplexSupp", "PositivityFct", "NegativityFct", "L1Norms", "L1NormsConj", "SSD", "L1Data", "L1DataConj",
           "QuadEpiInd", "QuadEpiSupp", "QuadSupport", "HuberPerspective", "TruncQuadEpiInd",
           "quad_dual_coefficients"]
