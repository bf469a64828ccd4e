"""Sums of per-point matrix norms (reference: opymize/functionals/l1norms.py).

Blocks of shape (M1, M2) are read column by column, consistent with
L1NormsProj; the Frobenius norm does not depend on that choice."""
import numpy as np

from opymize_amd import Functional, Variable, _hip
from opymize_amd.functionals._base import fn_eval
from opymize_amd.operators.proj import L1NormsProj


def _norm_params(M, matrixnorm, lbd):
    prm = _hip.FnParams()
    prm.m1, prm.m2 = int(M[0]), int(M[1])
    prm.matrixnorm = ord("F") if matrixnorm == "frobenius" else ord("N")
    prm.lbd = float(lbd)
    return prm


class L1Norms(Functional):
    """F(x) = lbd * sum_i |x_i| for the Frobenius or nuclear norm
    (reference: l1norms.py:34-70)."""

    def __init__(self, N, M, lbd, matrixnorm="frobenius", conj=None):
        Functional.__init__(self)
        assert matrixnorm in ["frobenius", "nuclear"]
        self.N, self.M = int(N), tuple(int(m) for m in M)
        self.x = Variable((self.N,) + self.M)
        self.lbd = float(lbd)
        self.matrixnorm = matrixnorm
        conjnorm = "spectral" if matrixnorm == "nuclear" else "frobenius"
        self.conj = L1NormsConj(N, M, lbd, conjnorm, conj=self) if conj is None else conj

    def _gpu_eval(self, x, out):
        fn_eval(_hip.FN_NORMS, x, self.M[0] * self.M[1], out, _norm_params(self.M, self.matrixnorm, self.lbd))

    def _gpu_finish(self, v):
        return self.lbd * float(v[0]), 0.0


class L1NormsConj(Functional):
    """F(x) = sum_i indicator(|x_i| <= lbd) for the Frobenius or spectral norm;
    prox = L1NormsProj, independent of the step (reference: l1norms.py:72-92)."""
    device_step = True  # prox accepts a step held in device memory


    def __init__(self, N, M, lbd, matrixnorm="frobenius", conj=None):
        Functional.__init__(self)
        assert matrixnorm in ["frobenius", "spectral"]
        self.N, self.M = int(N), tuple(int(m) for m in M)
        self.x = Variable((self.N,) + self.M)
        self.lbd = float(lbd)
        self.matrixnorm = matrixnorm
        conjnorm = "nuclear" if matrixnorm == "spectral" else "frobenius"
        self.conj = L1Norms(N, M, lbd, conjnorm, conj=self) if conj is None else conj
        self._prox = L1NormsProj(N, M, self.lbd, matrixnorm)

    def _gpu_eval(self, x, out):
        mn = "frobenius" if self.matrixnorm == "frobenius" else "nuclear"
        fn_eval(_hip.FN_NORMS, x, self.M[0] * self.M[1], out, _norm_params(self.M, mn, self.lbd))

    def _gpu_finish(self, v):
        return 0.0, float(v[1])

    def prox(self, tau):
        return self._prox
