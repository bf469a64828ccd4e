"""Per-coordinate loop versions of the finite-difference operators, written
from the reference's kernel semantics (diff.cu:37-243) independently of the
sparse builders in ops.py; used on small grids to pin them.
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py)."""
import itertools

import numpy as np


def grad_loops(x, dims, C, h, w, scheme="forward"):
    dims = tuple(dims)
    D = len(dims)
    X = x.reshape(dims + (C,))
    Y = np.zeros(dims + (D, C))
    for c in itertools.product(*[range(d) for d in dims]):
        on_br = any(c[a] == dims[a] - 1 for a in range(D))
        for t in range(D):
            for k in range(C):
                if scheme == "forward":
                    if c[t] < dims[t] - 1:
                        cp = list(c)
                        cp[t] += 1
                        Y[c + (t, k)] = w[k] / h[t] * (X[tuple(cp) + (k,)] - X[c + (k,)])
                else:
                    if on_br:
                        continue
                    acc = 0.0
                    others = [a for a in range(D) if a != t]
                    for bits in itertools.product([0, 1], repeat=len(others)):
                        base = list(c)
                        for a, b in zip(others, bits):
                            base[a] += b
                        up = list(base)
                        up[t] += 1
                        acc += X[tuple(up) + (k,)] - X[tuple(base) + (k,)]
                    Y[c + (t, k)] = w[k] / (h[t] * 2 ** (D - 1)) * acc
    return Y.ravel()


def div_loops(p, dims, C, h, w):
    """forward-scheme negative divergence (diff.cu:114-123)."""
    dims = tuple(dims)
    D = len(dims)
    P = p.reshape(dims + (D, C))
    X = np.zeros(dims + (C,))
    for c in itertools.product(*[range(d) for d in dims]):
        for k in range(C):
            v = 0.0
            for t in range(D):
                if c[t] < dims[t] - 1:
                    v -= P[c + (t, k)] / h[t]
                if c[t] > 0:
                    cm = list(c)
                    cm[t] -= 1
                    v += P[tuple(cm) + (t, k)] / h[t]
            X[c + (k,)] = w[k] * v
    return X.ravel()


def laplacian_loops(x, dims, C, h, bc):
    """Laplacian with the boundary rules of diff.cu:131-243."""
    dims = tuple(dims)
    D = len(dims)
    adj = bc.endswith("_adj")
    b = bc[0].upper()
    X = x.reshape(dims + (C,))
    Y = np.zeros_like(X)

    def at(c, t, s):
        cc = list(c)
        cc[t] += s
        return X[tuple(cc)]

    for c in itertools.product(*[range(d) for d in dims]):
        xi = X[c]
        v = np.zeros(C)
        if b == "C":
            bnd = [t for t in range(D) if c[t] == 0 or c[t] == dims[t] - 1]
            if adj:
                if len(bnd) >= 2:
                    pass
                elif len(bnd) == 1:
                    t = bnd[0]
                    ih2 = 1.0 / h[t] ** 2
                    v += ih2 * (at(c, t, 1) if c[t] == 0 else at(c, t, -1))
                else:
                    for t in range(D):
                        ih2 = 1.0 / h[t] ** 2
                        v -= 2 * ih2 * xi
                        if c[t] - 1 > 0:
                            v += ih2 * at(c, t, -1)
                        if c[t] + 1 < dims[t] - 1:
                            v += ih2 * at(c, t, 1)
            elif not bnd:
                for t in range(D):
                    ih2 = 1.0 / h[t] ** 2
                    v += ih2 * (at(c, t, -1) - xi) + ih2 * (at(c, t, 1) - xi)
        elif b == "S":
            for t in range(D):
                ih2 = 1.0 / h[t] ** 2
                inner = 0 < c[t] < dims[t] - 1
                if adj:
                    if inner:
                        v -= 2 * ih2 * xi
                    if c[t] > 1:
                        v += ih2 * at(c, t, -1)
                    if c[t] < dims[t] - 2:
                        v += ih2 * at(c, t, 1)
                elif inner:
                    v += ih2 * (at(c, t, -1) - 2 * xi + at(c, t, 1))
        else:
            for t in range(D):
                ih2 = 1.0 / h[t] ** 2
                if c[t] > 0:
                    v += ih2 * (at(c, t, -1) - xi)
                if c[t] < dims[t] - 1:
                    v += ih2 * (at(c, t, 1) - xi)
        Y[c] = v
    return Y.ravel()
