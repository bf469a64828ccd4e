"""Exact 1-D total-variation denoising, an independent pin for PDHG.
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

argmin_x 0.5*|x - f|^2 + mu*sum_i |x_{i+1} - x_i| has the dual
min_{|u|_inf <= mu} 0.5*|D^T u - f|^2 with x = f - D^T u, a bound-constrained
least-squares problem solved to machine precision by scipy's BVLS."""
import numpy as np
import scipy.sparse as sp
from scipy.optimize import lsq_linear


def tv1d_denoise(f, mu):
    f = np.asarray(f, dtype=np.float64).ravel()
    n = f.size
    if n == 1:
        return f.copy(), np.zeros(0)
    Dm = sp.diags([-np.ones(n - 1), np.ones(n - 1)], [0, 1], shape=(n - 1, n)).toarray()
    res = lsq_linear(Dm.T, f, bounds=(-mu, mu), method="bvls", tol=1e-15, max_iter=10 * n + 100)
    u = res.x
    return f - Dm.T.dot(u), u
