"""Solvers (reference: opymize/solvers/__init__.py)."""
from opymize_amd.solvers.pdhg import PDHG

__all__ = ["PDHG"]
