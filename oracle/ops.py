"""Sparse-matrix restatement of the hot path's linear operators (FP64).
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Layouts follow the reference: an image is (N, C), a vector field (N, D, C),
flat in C order (reference opymize/linear/diff.py:117-118)."""
from functools import reduce

import numpy as np
import scipy.sparse as sp


def skips(dims):
    s = np.ones(len(dims), dtype=np.int64)
    for t in range(len(dims) - 2, -1, -1):
        s[t] = s[t + 1] * dims[t + 1]
    return s


def _kron_all(mats):
    return reduce(lambda a, b: sp.kron(a, b, format="csr"), mats)


def _fwd1d(n):
    """forward difference with Neumann end: row i = e_{i+1} - e_i, last row 0
    (reference sparse.py:122-126)."""
    d = sp.lil_matrix((n, n))
    for i in range(n - 1):
        d[i, i] = -1.0
        d[i, i + 1] = 1.0
    return d.tocsr()


def _avg1d(n):
    """0.5*(e_i + e_{i+1}) with the last row zero (cells of the centered scheme)."""
    a = sp.lil_matrix((n, n))
    for i in range(n - 1):
        a[i, i] = 0.5
        a[i, i + 1] = 0.5
    return a.tocsr()


def _interleave(parts, N, D):
    """Stack per-axis point operators P_t (N x n) so that row i*D + t = P_t[i]."""
    out = None
    for t, P in enumerate(parts):
        e = sp.csr_matrix(([1.0], ([t], [0])), shape=(D, 1))
        term = sp.kron(P, e, format="csr")
        out = term if out is None else out + term
    return out.tocsr()


def grad_matrix(dims, C=1, h=None, w=None, scheme="forward"):
    """Matrix of GradientOp: (N*D*C) x (N*C)
    (reference diff.py:67-144; forward: diff.cu:64-71, centered: diff.cu:53-63)."""
    dims = [int(d) for d in dims]
    D = len(dims)
    N = int(np.prod(dims))
    h = np.ones(D) if h is None else np.asarray(h, dtype=np.float64)
    w = np.ones(C) if w is None else np.asarray(w, dtype=np.float64)
    parts = []
    for t in range(D):
        factors = []
        for a in range(D):
            if a == t:
                factors.append(_fwd1d(dims[a]) / h[t])
            elif scheme == "centered":
                factors.append(_avg1d(dims[a]))
            else:
                factors.append(sp.identity(dims[a], format="csr"))
        parts.append(_kron_all(factors))
    P = _interleave(parts, N, D)
    return sp.kron(P, sp.diags(w), format="csr")


def div_matrix(dims, C=1, h=None, w=None, scheme="forward"):
    """DivergenceOp = adjoint of GradientOp (reference diff.py:170)."""
    return grad_matrix(dims, C, h, w, scheme).T.tocsr()


def laplacian_matrix(dims, C=1, h=None, bc="neumann"):
    """LaplacianOp (reference diff.py:185-249 and sparse.py:11-62): sum over
    axes of 1-D second differences; 'neumann' uses one-sided ends,
    'second-order' zero rows at each axis' ends, 'curvature' additionally
    zeroes every row of a boundary point; '_adj' is the transpose."""
    dims = [int(d) for d in dims]
    D = len(dims)
    h = np.ones(D) if h is None else np.asarray(h, dtype=np.float64)
    base = bc[:-4] if bc.endswith("_adj") else bc

    def dd(n):
        m = sp.lil_matrix((n, n))
        for i in range(n):
            if base == "neumann":
                if i > 0:
                    m[i, i - 1] += 1.0
                    m[i, i] -= 1.0
                if i < n - 1:
                    m[i, i + 1] += 1.0
                    m[i, i] -= 1.0
            else:  # curvature / second-order: interior rows only
                if 0 < i < n - 1:
                    m[i, i - 1] = 1.0
                    m[i, i] = -2.0
                    m[i, i + 1] = 1.0
        return m.tocsr()

    N = int(np.prod(dims))
    L = sp.csr_matrix((N, N))
    for t in range(D):
        f = [sp.identity(dims[a], format="csr") for a in range(D)]
        f[t] = dd(dims[t]) / h[t] ** 2
        L = L + _kron_all(f)
    if base == "curvature":
        keep = []
        for n in dims:
            v = np.ones(n)
            v[[0, -1]] = 0.0
            keep.append(sp.diags(v))
        L = _kron_all(keep).dot(L)
    M = sp.kron(L, sp.identity(C), format="csr")
    return (M.T if bc.endswith("_adj") else M).tocsr()


def rowwise_lp(M, p=1):
    """sum_j |M_ij|^p per row (reference LinOp.rowwise_lp, __init__.py:122-142;
    identical to the reference's row norm for p = 1)."""
    A = abs(M.tocsr())
    A.data = A.data ** p
    return np.asarray(A.sum(axis=1)).ravel()


def matmul(A, x, trans=False):
    """MatrixMult: y_ji = sum_k A_jk x_ki (trans: A_kj) (reference einsum.py:17-56)."""
    return np.einsum("kj,ki->ji" if trans else "jk,ki->ji", A, x)


def matmul_r(A, x, trans=False):
    """MatrixMultR: y_ij = sum_k x_ik A_kj (trans: A_jk) (reference einsum.py:58-97)."""
    return np.einsum("jk,ik->ij" if trans else "kj,ik->ij", A, x)


def matmul_matrix(A, N, trans=False):
    """MatrixMult as a sparse matrix acting on flat (K, N) -> (J, N)."""
    Am = A.T if trans else A
    return sp.kron(sp.csr_matrix(Am), sp.identity(N), format="csr")


def matmul_r_matrix(A, N, trans=False):
    """MatrixMultR as a sparse matrix acting on flat (N, K) -> (N, J)."""
    Am = A.T if trans else A  # (K, J)
    return sp.kron(sp.identity(N), sp.csr_matrix(Am.T), format="csr")


# ---- operators of lifting models (reference linear/indexed.py, einsum.py:130-214) ----

def indexedmult(B, P, x):
    """IndexedMult: y[j, i, m] = -sum_l B[j, m, l] x[i, P[j, l]] with x (N, K),
    B (J, M, L), P (J, L) -> y (J, N, M) (reference indexed.py:63-94, docstring :64; indexed.cu:27-48)."""
    xg = x[:, P]  # (N, J, L): x[i, P[j, l]]
    return -np.einsum("jml,ijl->jim", B, xg)


def indexedmult_adj(B, P, x, K):
    """IndexedMultAdj: y[i, P[j, l]] -= sum_m B[j, m, l] x[j, i, m] with
    x (J, N, M) -> y (N, K) (reference indexed.py:34-61, docstring :35; indexed.cu:1-25)."""
    J, N, M = x.shape
    y = np.zeros((N, K))
    contrib = np.einsum("jml,jim->jil", B, x)  # (J, N, L)
    for j in range(J):
        for l in range(P.shape[1]):
            y[:, P[j, l]] -= contrib[j, :, l]
    return y


def tangled_matmul_r(A, x, trans=False):
    """TangledMatrixMultR: y_mik = sum_jl x_jil A_jlmk, trans: A_mkjl
    (reference einsum.py:130-173, docstring :131-132, the subscripts of its spmat :146)."""
    return np.einsum("mkjl,jil->mik" if trans else "jlmk,jil->mik", A, x)


def matmul_r_batched(A, x, trans=False):
    """MatrixMultRBatched: y_jlk = sum_m x_jlm A_jmk, trans: A_jkm
    (reference einsum.py:175-214, docstring :176-177, the subscripts of its spmat :187)."""
    return np.einsum("jkm,jlm->jlk" if trans else "jmk,jlm->jlk", A, x)


def indexedmult_matrix(B, P, K, N):
    """IndexedMult as a sparse matrix on flat x (N, K) -> y (J, N, M), entry by
    entry from the definition (duplicate indices in P add up)."""
    J, M, L = B.shape
    rows, cols, vals = [], [], []
    for j in range(J):
        for i in range(N):
            for m in range(M):
                for l in range(L):
                    rows.append((j * N + i) * M + m)
                    cols.append(i * K + P[j, l])
                    vals.append(-B[j, m, l])
    return sp.csr_matrix((vals, (rows, cols)), shape=(J * N * M, N * K))


def tangled_matmul_r_matrix(A, N, trans=False):
    """TangledMatrixMultR as a sparse matrix on flat x (J, N, L) -> y (M, N, K),
    entry by entry from the definition."""
    if trans:
        M, K, J, L = A.shape
        a = lambda j, l, m, k: A[m, k, j, l]
    else:
        J, L, M, K = A.shape
        a = lambda j, l, m, k: A[j, l, m, k]
    rows, cols, vals = [], [], []
    for m in range(M):
        for i in range(N):
            for k in range(K):
                for j in range(J):
                    for l in range(L):
                        rows.append((m * N + i) * K + k)
                        cols.append((j * N + i) * L + l)
                        vals.append(a(j, l, m, k))
    return sp.csr_matrix((vals, (rows, cols)), shape=(M * N * K, J * N * L))


def matmul_r_batched_matrix(A, N, trans=False):
    """MatrixMultRBatched as a sparse matrix on flat x (J, N, M) -> y (J, N, K),
    entry by entry from the definition."""
    if trans:
        J, K, M = A.shape
        a = lambda j, m, k: A[j, k, m]
    else:
        J, M, K = A.shape
        a = lambda j, m, k: A[j, m, k]
    rows, cols, vals = [], [], []
    for j in range(J):
        for i in range(N):
            for k in range(K):
                for m in range(M):
                    rows.append((j * N + i) * K + k)
                    cols.append((j * N + i) * M + m)
                    vals.append(a(j, m, k))
    return sp.csr_matrix((vals, (rows, cols)), shape=(J * N * K, J * N * M))
