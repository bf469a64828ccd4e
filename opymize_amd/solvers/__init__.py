"""Solvers (reference: opymize/solvers/__init__.py)."""
from opymize_amd.solvers.pdhg import PDHG
from opymize_amd.solvers.semismooth_newton import SemismoothNewton

__all__ = ["PDHG", "SemismoothNewton"]
