"""Helpers for evaluating functionals with the FP64 reduction kernel."""
import numpy as np

from opymize_amd import _hip


def fn_eval(kind, x, width, out, prm=None, accumulate=False):
    """Enqueue one FP64 reduction of `kind` over x (npoints x width) into the
    device double[3] `out` = (sum of values, max infeas 1, max infeas 2)."""
    if prm is None:
        prm = _hip.FnParams()
    n = x.numel()
    if n % width:
        raise ValueError("array of %d entries is not a multiple of %d" % (n, width))
    _hip.call("fn_eval", x.dtype, kind, _hip.ptr(x), n // width, int(width), _hip.ctypes.byref(prm),
              _hip.ptr(_hip.Workspace.get()), _hip.ptr(out), int(accumulate), _hip.stream())


def vptr(t):
    return None if t is None else t.data_ptr()


class Slots(object):
    """Mixin: number of (sum, max, max) result triples a functional needs."""

    def nslots(self):
        return 1
