"""Solvers (reference: opymize/solvers/__init__.py)."""
from opymize_amd.solvers.pdhg import PDHG
from opymize_amd.solvers.semismooth_newton import SemismoothNewton, SemismoothQuasinewton

__all__ = ["PDHG", "SemismoothNewton", "SemismoothQuasinewton"]
