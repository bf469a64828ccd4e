"""Operator test helpers (reference: opymize/tools/tests.py:47-95), running the
operators on the GPU."""
import numpy as np


def test_adjoint(op, dtype=np.float64, tol_per_entry=1e-11):
    """<Ax, y> == <x, A^T y> within N*tol (reference: tools/tests.py:47-56)."""
    N = max(op.x.size, op.y.size)
    x = np.random.randn(op.x.size).astype(dtype)
    y = np.random.randn(op.y.size).astype(dtype)
    ATy = np.zeros_like(x)
    Ax = np.zeros_like(y)
    op(x, Ax)
    op.adjoint(y, ATy)
    err = abs(float(Ax.astype(np.float64).dot(y)) - float(x.astype(np.float64).dot(ATy)))
    assert err < N * tol_per_entry, "adjoint mismatch %g" % err
    return err


def test_rowwise_lp(op, p=1, maxiter=10, dtype=np.float64):
    """row-wise l^p norms against explicit rows A^T e_i (reference: tools/tests.py:58-67)."""
    for _ in range(maxiter):
        i = np.random.randint(0, op.y.size)
        y = op.y.new(dtype=dtype)
        y[i] = 1.0
        Ai = op.x.new(dtype=dtype)
        op.adjoint(y, Ai)
        op.rowwise_lp(y, p=p)
        assert abs(y[i] - np.sum(np.abs(Ai) ** p)) < 1e-10 * max(1.0, abs(y[i]))
