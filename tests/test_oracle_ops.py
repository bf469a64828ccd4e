"""Pins of the oracle's linear operators (CPU only)."""
import numpy as np
import pytest

from oracle import ops, ops_loops


@pytest.mark.parametrize("dims,C", [((4, 5), 1), ((3, 4), 3), ((2, 3, 4), 2), ((7,), 2)])
@pytest.mark.parametrize("scheme", ["forward", "centered"])
def test_grad_matrix_matches_loops(dims, C, scheme):
    rng = np.random.default_rng(0)
    h = rng.uniform(0.5, 2.0, len(dims))
    w = rng.uniform(0.5, 2.0, C)
    x = rng.standard_normal(int(np.prod(dims)) * C)
    G = ops.grad_matrix(dims, C, h, w, scheme)
    np.testing.assert_allclose(G.dot(x), ops_loops.grad_loops(x, dims, C, h, w, scheme), rtol=0, atol=1e-13)


@pytest.mark.parametrize("dims,C", [((4, 5), 1), ((3, 4), 3), ((2, 3, 4), 2)])
def test_div_is_adjoint_and_matches_loops(dims, C):
    rng = np.random.default_rng(1)
    h = rng.uniform(0.5, 2.0, len(dims))
    w = rng.uniform(0.5, 2.0, C)
    N = int(np.prod(dims))
    p = rng.standard_normal(N * len(dims) * C)
    Dv = ops.div_matrix(dims, C, h, w)
    np.testing.assert_allclose(Dv.dot(p), ops_loops.div_loops(p, dims, C, h, w), rtol=0, atol=1e-13)
    x = rng.standard_normal(N * C)
    G = ops.grad_matrix(dims, C, h, w)
    assert abs(G.dot(x).dot(p) - x.dot(Dv.dot(p))) < 1e-11 * N


@pytest.mark.parametrize("bc", ["neumann", "curvature", "curvature_adj", "second-order", "second-order_adj"])
def test_laplacian_matrix_matches_loops(bc):
    rng = np.random.default_rng(2)
    dims, C = (5, 6), 2
    h = np.array([0.7, 1.3])
    x = rng.standard_normal(30 * C)
    L = ops.laplacian_matrix(dims, C, h, bc)
    np.testing.assert_allclose(L.dot(x), ops_loops.laplacian_loops(x, dims, C, h, bc), rtol=0, atol=1e-12)


def test_laplacian_adjoint_pairs():
    rng = np.random.default_rng(3)
    dims = (6, 7)
    for bc in ["curvature", "second-order"]:
        L = ops.laplacian_matrix(dims, 1, None, bc)
        La = ops.laplacian_matrix(dims, 1, None, bc + "_adj")
        x, y = rng.standard_normal(42), rng.standard_normal(42)
        assert abs(L.dot(x).dot(y) - x.dot(La.dot(y))) < 1e-11


# exact table of the reference's tests/sparse.py:59-93 (centered scheme, 4x5 grid)
TEST_V = np.array([[0., 0., 0., 0., 0.],
                   [0., 1., 0., 0., 0.],
                   [0., 0., 0., 0., 1.],
                   [0., 0., 0., 0., 0.]])
TEST_DV_CENTERED = np.array([[[0.5, 0.5, 0., 0., 0.],
                              [-0.5, -0.5, 0., 0.5, 0.],
                              [0., 0., 0., -0.5, 0.],
                              [0., 0., 0., 0., 0.]],
                             [[0.5, -0.5, 0., 0., 0.],
                              [0.5, -0.5, 0., 0.5, 0.],
                              [0., 0., 0., 0.5, 0.],
                              [0., 0., 0., 0., 0.]]])


@pytest.mark.parametrize("c", [1, 2, 3])
def test_centered_scheme_reference_table(c):
    cc = min(1, c - 1)
    Dv = TEST_DV_CENTERED.transpose(1, 2, 0).ravel()
    rng = np.random.default_rng(4)
    w = rng.standard_normal(c * TEST_V.size)
    w.reshape(-1, c)[:, cc] = TEST_V.ravel()
    Dw = ops.grad_matrix(TEST_V.shape, c, scheme="centered").dot(w)
    assert np.linalg.norm(Dw.reshape(-1, c)[:, cc] - Dv, ord=np.inf) == 0.0


def _cell_grid(shape):
    dom = np.array([[-np.pi, np.pi], [-np.pi, np.pi]])
    h = (dom[:, 1] - dom[:, 0]) / np.array(shape)
    g = np.mgrid[[slice(0.0, s) for s in shape]].reshape(2, -1).T * h[None, :] + dom[None, :, 0] + 0.5 * h
    return g, h


@pytest.mark.parametrize("s", [1, 2])
def test_grad_converges_to_analytic(s):
    # reference tests/diff.py:42-57 (with the forward scheme the derivative
    # sits at the cell edge midpoints, so compare at the half-shifted grid per axis)
    shape = (10 ** s, 10 ** s)
    grid, h = _cell_grid(shape)
    fx = np.sin(grid[:, 0]) + grid[:, 0] ** 2 * np.cos(grid[:, 1])
    y = ops.grad_matrix(shape, 1, h, scheme="centered").dot(fx).reshape(shape + (2,))
    g2 = grid + 0.5 * h[None, :]
    gx = np.cos(g2[:, 0]) + 2 * g2[:, 0] * np.cos(g2[:, 1])
    gy = -g2[:, 0] ** 2 * np.sin(g2[:, 1])
    ytest = np.stack([gx, gy], axis=-1).reshape(shape + (2,))
    vol = np.prod(h)
    assert np.linalg.norm(vol * (y[:-1, :-1] - ytest[:-1, :-1]).ravel()) < 2 * vol


@pytest.mark.parametrize("bc", ["curvature", "neumann", "second-order"])
def test_laplacian_converges_to_analytic(bc):
    # reference tests/diff.py:69-82
    shape = (100, 100)
    grid, h = _cell_grid(shape)
    fx = np.sin(grid[:, 0]) + grid[:, 0] ** 2 * np.cos(grid[:, 1])
    y = ops.laplacian_matrix(shape, 1, h, bc).dot(fx).reshape(shape)
    lap = ((2 - grid[:, 0] ** 2) * np.cos(grid[:, 1]) - np.sin(grid[:, 0])).reshape(shape)
    vol = np.prod(h)
    assert np.linalg.norm(vol * (y[1:-1, 1:-1] - lap[1:-1, 1:-1]).ravel()) < 2 * vol


def test_curvature_laplacian_ghost_cells():
    # reference tests/diff.py:84-107: the curvature Laplacian equals the
    # Laplacian of the ghost-cell extension whose interior Laplacian is
    # minimal in least squares (cvxpy there, a sparse lstsq here).
    import scipy.sparse as sp
    from scipy.sparse.linalg import lsqr
    rng = np.random.default_rng(5)
    m, n = 12, 15
    data = rng.random((m, n))
    Dy_curv = ops.laplacian_matrix((m, n), 1, None, "curvature").dot(data.ravel()).reshape(m, n)
    M, N = m + 2, n + 2
    dd = lambda k: sp.diags([np.ones(k - 1), -2 * np.ones(k), np.ones(k - 1)], [-1, 0, 1])
    L = (sp.kron(dd(M), sp.identity(N)) + sp.kron(sp.identity(M), dd(N))).tocsr()
    inner = np.zeros((M, N), dtype=bool)
    inner[1:-1, 1:-1] = True
    inner = inner.ravel()
    rows = np.flatnonzero(inner)
    z = np.zeros(M * N)
    z[inner] = data.ravel()
    Lg = L[rows][:, ~inner]
    rhs = -L[rows][:, inner].dot(data.ravel())
    g = lsqr(Lg, rhs, atol=1e-15, btol=1e-15, iter_lim=100000)[0]
    z[~inner] = g
    Dy = L.dot(z)[rows].reshape(m, n)
    # the ghost-cell solution is not unique at corners, compare away from them
    assert np.abs(Dy_curv[1:-1, 1:-1] - Dy[1:-1, 1:-1]).max() < 1e-9


EINSUM_CASES = [((5, 7), 11, False), ((5, 7), 11, True), ((16, 16), 32, False), ((3, 30), 4, True)]


@pytest.mark.parametrize("shape,N,trans", EINSUM_CASES)
def test_matmul_matrices_match_einsum(shape, N, trans):
    rng = np.random.default_rng(6)
    A = rng.standard_normal(shape)
    K = shape[0] if trans else shape[1]
    x = rng.standard_normal((K, N))
    np.testing.assert_allclose(ops.matmul_matrix(A, N, trans).dot(x.ravel()), ops.matmul(A, x, trans).ravel(),
                               atol=1e-13)
    Kr = shape[1] if trans else shape[0]
    xr = rng.standard_normal((N, Kr))
    np.testing.assert_allclose(ops.matmul_r_matrix(A, N, trans).dot(xr.ravel()), ops.matmul_r(A, xr, trans).ravel(),
                               atol=1e-13)
    # independent restatement: explicit per-point products
    yl = np.stack([(A.T if trans else A).dot(x[:, i]) for i in range(N)], axis=1)
    np.testing.assert_allclose(ops.matmul(A, x, trans), yl, atol=1e-13)
    yr = np.stack([xr[i].dot(A.T if trans else A) for i in range(N)])
    np.testing.assert_allclose(ops.matmul_r(A, xr, trans), yr, atol=1e-13)


def test_rowwise_lp_against_explicit_rows():
    # reference opymize/tools/tests.py:58-67
    G = ops.grad_matrix((6, 5), 3, np.array([0.3, 0.4]), np.array([1.0, 2.0, 0.5]))
    r = ops.rowwise_lp(G, 1)
    Gd = G.toarray()
    np.testing.assert_allclose(r, np.abs(Gd).sum(axis=1), atol=1e-12)


# ---- lifting operators (reference linear/indexed.py, einsum.py:130-214) ----------------

def _lifting_case(seed, J=4, M=2, L=3, K=6, N=5, dup=True):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((J, M, L))
    P = rng.integers(0, K, (J, L))
    if dup:
        P[0, 2] = P[0, 0]  # a repeated label index: its coefficients add up
    return B, P, K, N, rng


def test_indexedmult_matches_its_definition_and_adjoint():
    """einsum restatement = the matrix built entry by entry from the
    docstring formula (indexed.py:64); the scatter restatement of the adjoint
    (indexed.py:35) = that matrix's transpose."""
    B, P, K, N, rng = _lifting_case(0)
    Mx = ops.indexedmult_matrix(B, P, K, N)
    x = rng.standard_normal((N, K))
    np.testing.assert_allclose(ops.indexedmult(B, P, x).ravel(), Mx @ x.ravel(), atol=1e-14)
    z = rng.standard_normal((B.shape[0], N, B.shape[1]))
    np.testing.assert_allclose(ops.indexedmult_adj(B, P, z, K).ravel(), Mx.T @ z.ravel(), atol=1e-14)
    # a hand example: one region, one output, labels 2 and 0 with weights 3 and -1
    Bh = np.array([[[3.0, -1.0]]])
    Ph = np.array([[2, 0]])
    xh = np.array([[10.0, 20.0, 30.0]])
    np.testing.assert_allclose(ops.indexedmult(Bh, Ph, xh), [[[-(3 * 30.0 - 10.0)]]])
    np.testing.assert_allclose(ops.indexedmult_adj(Bh, Ph, np.array([[[2.0]]]), 3), [[2.0, 0.0, -6.0]])


@pytest.mark.parametrize("trans", [False, True])
def test_tangled_and_batched_match_their_definitions(trans):
    rng = np.random.default_rng(1)
    N = 7
    A4 = rng.standard_normal((2, 3, 4, 5))
    x = rng.standard_normal(((4, N, 5) if trans else (2, N, 3)))
    Mt = ops.tangled_matmul_r_matrix(A4, N, trans)
    np.testing.assert_allclose(ops.tangled_matmul_r(A4, x, trans).ravel(), Mt @ x.ravel(), atol=1e-13)
    A3 = rng.standard_normal((3, 4, 2))
    xb = rng.standard_normal(((3, N, 2) if trans else (3, N, 4)))
    Mb = ops.matmul_r_batched_matrix(A3, N, trans)
    np.testing.assert_allclose(ops.matmul_r_batched(A3, xb, trans).ravel(), Mb @ xb.ravel(), atol=1e-13)
    # the transposed operator is the adjoint of the plain one
    Mt0 = ops.tangled_matmul_r_matrix(A4, N, not trans)
    np.testing.assert_allclose((Mt - Mt0.T).toarray(), 0.0, atol=0)
    Mb0 = ops.matmul_r_batched_matrix(A3, N, not trans)
    np.testing.assert_allclose((Mb - Mb0.T).toarray(), 0.0, atol=0)
    # tangled = an ordinary product after grouping (i)(jl) x (jl)(mk) (einsum.py:133-135)
    if not trans:
        J, L, M, K = A4.shape
        X = x.transpose(1, 0, 2).reshape(N, J * L)
        Y = X @ A4.reshape(J * L, M * K)
        np.testing.assert_allclose(ops.tangled_matmul_r(A4, x).transpose(1, 0, 2).reshape(N, M * K), Y, atol=1e-13)
