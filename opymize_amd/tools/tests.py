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


def compare_vars(descr, v1, v2, rtol=1e-5, atol=1e-8):
    """Compare two flat arrays block by block of the Variable `descr`; raise
    on the first block that differs (reference: tools/tests.py:90-95)."""
    for i, (a, b) in enumerate(zip(descr.vars(np.asarray(v1)), descr.vars(np.asarray(v2)))):
        if not np.allclose(a, b, rtol=rtol, atol=atol):
            raise AssertionError("block %d differs: max |a - b| = %e" % (i, np.abs(a - b).max()))


def test_gpu_op(op, reference, type_t="double", rtol=None, atol=None, seed=None):
    """Bootstrap test of an operator's GPU implementation (reference:
    tools/tests.py:69-88).  The reference compares its CPU and GPU paths;
    this package has only the GPU path, so the caller supplies the CPU
    result as `reference(x, y0) -> y` (x, y0 float64 host arrays; y0 is the
    output's prior content, which add=True accumulates onto).  Checks the
    overwrite and the add=True forms of op(x, y) on device arrays."""
    import torch
    np_dtype = np.float64 if type_t == "double" else np.float32
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(op.x.size).astype(np_dtype)
    y0 = rng.standard_normal(op.y.size).astype(np_dtype)
    if rtol is None:
        rtol = 1e-12 if type_t == "double" else 2e-6
    if atol is None:
        atol = 1e-12 if type_t == "double" else 2e-6
    x64, y064 = x.astype(np.float64), y0.astype(np.float64)
    want = np.asarray(reference(x64.copy(), y064.copy()), dtype=np.float64).ravel()
    for add in (False, True):
        xd = torch.as_tensor(x).cuda()
        yd = torch.as_tensor(y0).cuda()
        op(xd, yd, add=add)
        got = yd.cpu().numpy().astype(np.float64)
        compare_vars(op.y, got, want + (y064 if add else 0.0), rtol=rtol, atol=atol)
