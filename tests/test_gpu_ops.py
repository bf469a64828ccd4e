"""GPU parity of the operator / prox / reduction kernels against the oracle.

Tolerances: FP64 kernels agree with the FP64 oracle to 1e-12 relative
(rounding-order differences only: FMA contraction, summation order); FP32
kernels to 2e-5 relative to the data scale."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ops, prox

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 2e-5}


def dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


def host(t):
    return t.cpu().numpy().astype(np.float64)


def close(a, b, dtype, scale=None):
    scale = max(1.0, np.abs(b).max()) if scale is None else scale
    np.testing.assert_allclose(a, b, rtol=0, atol=TOL[dtype] * scale * 10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,C", [((7, 9), 1), ((16, 33), 3), ((5, 6, 7), 2), ((6, 11, 16), 1), ((300,), 1),
                                    ((1, 17), 2)])
@pytest.mark.parametrize("scheme", ["forward", "centered"])
def test_gradient_divergence(gpu, dtype, dims, C, scheme):
    from opymize_amd.linear import GradientOp
    rng = np.random.default_rng(0)
    h = rng.uniform(0.5, 2, len(dims))
    w = rng.uniform(0.5, 2, C)
    op = GradientOp(dims, C, scheme=scheme, imageh=h, weights=w)
    G = ops.grad_matrix(dims, C, h, w, scheme)
    x = rng.standard_normal(op.x.size)
    y0 = rng.standard_normal(op.y.size)
    xd, yd = dev(x, dtype), dev(y0, dtype)
    op(xd, yd)
    close(host(yd), G.dot(x.astype(dtype).astype(np.float64)), dtype)
    yd = dev(y0, dtype)
    op(xd, yd, add=True)
    close(host(yd), y0.astype(dtype) + G.dot(x.astype(dtype).astype(np.float64)), dtype)
    p = rng.standard_normal(op.y.size)
    xo = dev(np.zeros(op.x.size), dtype)
    op.adjoint(dev(p, dtype), xo)
    close(host(xo), G.T.dot(p.astype(dtype).astype(np.float64)), dtype)
    # row-wise l1 norms of both
    r = dev(np.zeros(op.y.size), dtype)
    op.rowwise_lp(r)
    close(host(r), ops.rowwise_lp(G, 1), dtype)
    r = dev(np.zeros(op.x.size), dtype)
    op.adjoint.rowwise_lp(r)
    close(host(r), ops.rowwise_lp(G.T, 1), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("C,offset", [(1, 1), (1, 0), (4, 0), (2, 3)])
def test_gradient_divergence_2d_views(gpu, dtype, C, offset):
    """The 2-D forward kernels on views at element offsets (the 2-vector
    accesses of C = 1 need an aligned field, else the scalar form runs), for
    C = 1, 2 and 4, with and without add, against the oracle matrices."""
    from opymize_amd.linear import GradientOp
    import torch
    rng = np.random.default_rng(7)
    dims = (37, 45)
    h = rng.uniform(0.5, 2, 2)
    w = rng.uniform(0.5, 2, C)
    op = GradientOp(dims, C, scheme="forward", imageh=h, weights=w)
    G = ops.grad_matrix(dims, C, h, w, "forward")
    x = rng.standard_normal(op.x.size)
    p = rng.standard_normal(op.y.size)
    xr = x.astype(dtype).astype(np.float64)
    pr = p.astype(dtype).astype(np.float64)
    for add in (False, True):
        yb = dev(np.concatenate([np.zeros(offset), rng.standard_normal(op.y.size)]), dtype)
        y0 = host(yb)[offset:]
        xb = dev(np.concatenate([np.zeros(offset), x]), dtype)
        op(xb[offset:], yb[offset:], add=add)
        close(host(yb)[offset:], (y0 if add else 0) + G.dot(xr), dtype)
        pb = dev(np.concatenate([np.zeros(offset), p]), dtype)
        ob = dev(np.concatenate([np.zeros(offset), rng.standard_normal(op.x.size)]), dtype)
        o0 = host(ob)[offset:]
        op.adjoint(pb[offset:], ob[offset:], add=add)
        close(host(ob)[offset:], (o0 if add else 0) + G.T.dot(pr), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bc", ["neumann", "curvature", "curvature_adj", "second-order", "second-order_adj"])
def test_laplacian(gpu, dtype, bc):
    from opymize_amd.linear import LaplacianOp
    dims, C = (10, 12), 3
    h = np.array([0.6, 1.4])
    op = LaplacianOp(dims, C, imageh=h, boundary=bc.replace("_adj", ""))
    if bc.endswith("_adj"):
        op = op.adjoint
    L = ops.laplacian_matrix(dims, C, h, bc)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(op.x.size)
    yd = dev(np.zeros(op.y.size), dtype)
    op(dev(x, dtype), yd)
    close(host(yd), L.dot(x.astype(dtype).astype(np.float64)), dtype)
    r = dev(np.zeros(op.y.size), dtype)
    op.rowwise_lp(r)
    close(host(r), ops.rowwise_lp(L, 1), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape,N,trans", [((5, 7), 33, False), ((5, 7), 33, True), ((16, 16), 1000, False),
                                           ((30, 15), 257, True), ((15, 30), 4099, False), ((15, 30), 4099, True),
                                           ((64, 64), 70, False), ((1, 1), 17, True), ((16, 33), 100, True),
                                           ((65, 3), 50, False), ((3, 65), 50, True)])
def test_matrix_mult(gpu, dtype, shape, N, trans):
    """Both layouts, both orientations; in FP32, K, J <= 64 run on the matrix
    cores (the 65-wide shapes take the scalar kernel)."""
    from opymize_amd.linear import MatrixMult, MatrixMultR
    rng = np.random.default_rng(2)
    A = rng.standard_normal(shape)
    for cls, mat in [(MatrixMult, ops.matmul_matrix), (MatrixMultR, ops.matmul_r_matrix)]:
        op = cls(N, A, trans=trans)
        M = mat(A, N, trans)
        x = rng.standard_normal(op.x.size)
        y0 = rng.standard_normal(op.y.size)
        yd = dev(y0, dtype)
        op(dev(x, dtype), yd, add=True)
        close(host(yd), y0.astype(dtype) + M.dot(x.astype(dtype).astype(np.float64)), dtype, scale=10)
        xo = dev(np.zeros(op.x.size), dtype)
        op.adjoint(dev(y0, dtype), xo)
        close(host(xo), M.T.dot(y0.astype(dtype).astype(np.float64)), dtype, scale=10)
        r = dev(np.zeros(op.y.size), dtype)
        op.rowwise_lp(r)
        close(host(r), ops.rowwise_lp(M, 1), dtype, scale=10)


@pytest.mark.parametrize("shape,trans", [((15, 30), False), ((16, 16), True), ((30, 15), False), ((7, 50), True)])
def test_matrix_mult_large(gpu, shape, trans):
    """FP32 on the matrix cores at 2^20 + 37 points (the staged kernels of
    both layouts, with a partial last round; K > J keeps the direct one)
    against numpy's einsum in FP64, plain and with add."""
    import torch
    from opymize_amd.linear import MatrixMult, MatrixMultR
    rng = np.random.default_rng(5)
    A = rng.standard_normal(shape)
    N = (1 << 20) + 37
    for cls in (MatrixMult, MatrixMultR):
        op = cls(N, A, trans=trans)
        x = rng.standard_normal(op.x.size).astype(np.float32)
        y0 = rng.standard_normal(op.y.size).astype(np.float32)
        if cls is MatrixMult:  # y[j, i] = sum_k A(j, k) x[k, i]
            Ajk = A.T if trans else A
            ref = np.einsum("jk,ki->ji", Ajk, x.reshape(-1, N).astype(np.float64)).ravel()
        else:  # y[i, j] = sum_k x[i, k] A(k, j)
            Akj = A.T if trans else A
            ref = np.einsum("ik,kj->ij", x.reshape(N, -1).astype(np.float64), Akj).ravel()
        for add in (False, True):
            yd = torch.as_tensor(y0.copy()).cuda()
            op(torch.as_tensor(x).cuda(), yd, add=add)
            want = ref + (y0.astype(np.float64) if add else 0.0)
            err = np.abs(yd.cpu().numpy().astype(np.float64) - want).max()
            assert err < 2e-5 * max(1.0, np.abs(want).max()), (cls.__name__, add, err)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_and_blockop(gpu, dtype):
    from opymize_amd.linear import BlockOp, CsrOp, GradientOp, ScaleOp, IdentityOp, DiagMatrixMultR
    rng = np.random.default_rng(3)
    M = sp.random(40, 30, density=0.2, random_state=4, format="csr")
    op = CsrOp(M)
    x = rng.standard_normal(30)
    yd = dev(np.zeros(40), dtype)
    op(dev(x, dtype), yd)
    close(host(yd), M.dot(x), dtype)
    G = GradientOp((5, 6), 1, scheme="forward")
    dm = rng.uniform(-1, 1, 2)
    B = BlockOp([[G, 0], [IdentityOp(30), ScaleOp(30, 2.5)], [0, DiagMatrixMultR(15, dm)]])
    Gm = ops.grad_matrix((5, 6), 1)
    Bm = sp.bmat([[Gm, None], [sp.identity(30), 2.5 * sp.identity(30)],
                  [None, sp.kron(sp.identity(15), sp.diags(dm))]]).tocsr()
    x = rng.standard_normal(B.x.size)
    yd = dev(rng.standard_normal(B.y.size), dtype)
    B(dev(x, dtype), yd)
    close(host(yd), Bm.dot(x), dtype)
    y = rng.standard_normal(B.y.size)
    xo = dev(np.zeros(B.x.size), dtype)
    B.adjoint(dev(y, dtype), xo)
    close(host(xo), Bm.T.dot(y), dtype)
    r = dev(np.zeros(B.y.size), dtype)
    B.rowwise_lp(r)
    close(host(r), ops.rowwise_lp(Bm, 1), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("M1,M2,norm", [(1, 2, "frobenius"), (3, 2, "frobenius"), (3, 2, "spectral"),
                                        (2, 3, "spectral"), (2, 2, "spectral"), (16, 2, "frobenius")])
def test_l1norms_proj(gpu, dtype, M1, M2, norm):
    from opymize_amd.operators import L1NormsProj
    rng = np.random.default_rng(5)
    N, lbd = 1000, 0.9
    x = rng.standard_normal(N * M1 * M2)
    op = L1NormsProj(N, (M1, M2), lbd, norm)
    xd = dev(x, dtype)
    op(xd)
    close(host(xd), prox.l1norms_proj(x.astype(dtype).astype(np.float64), N, M1, M2, lbd, norm), dtype)
    # out-of-place leaves the input alone
    xd, yd = dev(x, dtype), dev(np.zeros_like(x), dtype)
    op(xd, yd)
    close(host(xd), x.astype(dtype), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("M,lbd,shifted", [(1, np.inf, True), (3, np.inf, False), (2, 1.3, False),
                                           (1, 0.5, False), (4, np.inf, True)])
def test_quadepi_proj(gpu, dtype, M, lbd, shifted):
    from opymize_amd.operators import QuadEpiProj
    rng = np.random.default_rng(6)
    N, a = 2000, 1.7
    x = rng.standard_normal(N * (M + 1)) * 2
    shift = rng.standard_normal((N, M + 1)) if shifted else None
    op = QuadEpiProj(N, M, alph=a, shift=shift, lbd=lbd)
    xd = dev(x, dtype)
    op(xd)
    sh = None if shift is None else shift.astype(dtype).astype(np.float64)
    ref = prox.quadepi_proj(x.astype(dtype).astype(np.float64), N, M, a, sh, lbd)
    tol = 1e-11 if dtype == np.float64 else 2e-3  # FP32 cubic roots near the paraboloid's apex
    np.testing.assert_allclose(host(xd), ref, rtol=0, atol=tol * 10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ssd_and_l1data_prox(gpu, dtype):
    import torch
    from opymize_amd.functionals import SSD, L1Data
    rng = np.random.default_rng(7)
    N, C = 500, 3
    f = rng.standard_normal((N, C))
    vol = rng.uniform(0.5, 3, C)
    mask = rng.random(N) > 0.3
    x = rng.standard_normal(N * C)
    G = SSD(f, vol=vol, mask=mask)
    for tau in [0.37, rng.uniform(0.1, 1, N * C)]:
        xd = dev(x, dtype)
        t = tau if np.isscalar(tau) else dev(tau, dtype)
        G.prox(t)(xd)
        close(host(xd), prox.ssd_prox(x, f, vol, tau, mask), dtype)
    # step read from device memory
    st = torch.tensor([0.0, 0.81], dtype=torch.float64).cuda()
    xd = dev(x, dtype)
    G.prox((st, 1))(xd)
    close(host(xd), prox.ssd_prox(x, f, vol, 0.81, mask), dtype)
    # conjugate prox through the conjugate functional
    xd = dev(x, dtype)
    G.conj.prox(0.5)(xd)
    cm = np.zeros_like(f)
    cm[mask] = -f[mask] * vol
    close(host(xd), prox.ssd_prox(x, cm, 1 / vol, 0.5, mask), dtype)
    lbd = rng.uniform(0.5, 2, C)
    H = L1Data(f, lbd)
    xd = dev(x, dtype)
    H.prox(0.6)(xd)
    close(host(xd), prox.l1data_prox(x, f, lbd, 0.6), dtype)
    xd = dev(x, dtype)
    H.conj.prox(0.6)(xd)
    close(host(xd), prox.l1data_conj_prox(x, f, lbd, 0.6), dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_functional_values(gpu, dtype):
    from opymize_amd.functionals import (SSD, L1Norms, L1NormsConj, L1Data, QuadEpiInd, QuadEpiSupp,
                                         HuberPerspective, TruncQuadEpiInd, PositivityFct, IndicatorFct, SplitSum)
    rng = np.random.default_rng(8)
    N, C = 700, 3
    f = rng.standard_normal((N, C))
    vol = rng.uniform(0.5, 3, C)
    x = rng.standard_normal(N * C).astype(dtype).astype(np.float64)
    rt = 1e-10 if dtype == np.float64 else 1e-6
    v, i = SSD(f, vol=vol, shift=1.5)(dev(x, dtype))
    ff = f.astype(dtype).astype(np.float64)
    assert np.isclose(v, 0.5 * np.einsum("ik,k->", (x.reshape(N, C) - ff) ** 2, vol) + 1.5, rtol=rt)
    z = rng.standard_normal(N * 6).astype(dtype).astype(np.float64)
    for mn, sm, bm in [("frobenius", "frobenius", "frobenius"), ("nuclear", "nuclear", "spectral")]:
        F = L1Norms(N, (3, 2), 0.8, matrixnorm=mn)
        v, i = F(dev(z, dtype))
        assert np.isclose(v, 0.8 * prox.block_norms(z, N, 3, 2, sm).sum(), rtol=rt) and i == 0
        v, i = F.conj(dev(z, dtype))
        assert v == 0 and np.isclose(i, max(0, (prox.block_norms(z, N, 3, 2, bm) - 0.8).max()), rtol=rt)
    lbd = rng.uniform(0.5, 2, C)
    v, i = L1Data(f, lbd)(dev(x, dtype))
    assert np.isclose(v, (np.abs(x.reshape(N, C) - ff) * lbd).sum(), rtol=rt)
    v, i = L1Data(f, lbd).conj(dev(x, dtype))
    assert np.isclose(v, (x.reshape(N, C) * ff).sum(), rtol=rt)
    assert np.isclose(i, max(0, (np.abs(x.reshape(N, C)) - lbd).max()), rtol=rt)
    M = 2
    q = rng.standard_normal(N * (M + 1)).astype(dtype).astype(np.float64)
    qa, qb, qc = 1.3, rng.standard_normal((N, M)), rng.standard_normal(N)
    Q = q.reshape(N, M + 1)
    v, i = QuadEpiInd(N, M, a=qa, b=qb, c=qc)(dev(q, dtype))
    fx = (0.5 * qa * Q[:, :-1] ** 2 + qb * Q[:, :-1]).sum(1) + qc
    assert v == 0 and np.isclose(i, max(0, (fx - Q[:, -1]).max()), rtol=rt)
    v, i = QuadEpiSupp(N, M, a=qa, b=qb, c=qc)(dev(q, dtype))
    msk = Q[:, -1] < -1e-8
    x1, x2 = Q[msk, :-1], -Q[msk, -1]
    vv = (0.5 * qa * x1 ** 2 / x2[:, None] + qb[msk] * x1).sum(1) + x2 * qc[msk]
    assert np.isclose(v, vv.sum(), rtol=rt) and np.isclose(i, np.abs(Q[~msk, -1]).max(), rtol=rt)
    hl, ha = 1.7, 0.6
    v, i = HuberPerspective(N, M, lbd=hl, alph=ha)(dev(q, dtype))
    n1 = np.linalg.norm(Q[:, :-1], axis=1)
    x2 = -Q[:, -1]
    qm = n1 < ha * x2
    lm = (~qm) & (x2 > -1e-4)
    vv = hl * (0.5 / ha * (n1[qm] ** 2 / x2[qm]).sum() + (n1[lm] - x2[lm] * ha / 2).sum())
    assert np.isclose(v, vv, rtol=rt) and np.isclose(i, max(0, (-x2).max()), rtol=rt)
    v, i = TruncQuadEpiInd(N, M, lbd=hl, alph=ha)(dev(q, dtype))
    ii = max(0, (n1 - hl).max()) + max(0, (0.5 * ha * np.minimum(hl, n1) ** 2 - Q[:, -1]).max())
    assert v == 0 and np.isclose(i, ii, rtol=rt)
    S = SplitSum([PositivityFct(N), IndicatorFct(N * 2, c1=0.25, c2=3.0)])
    w = rng.standard_normal(3 * N).astype(dtype).astype(np.float64)
    v, i = S(dev(w, dtype))
    assert v == 3.0 and np.isclose(i, max(max(0, (-w[:N]).max()), np.abs(w[N:] - 0.25).max()), rtol=rt)


def test_normest_matches_oracle(gpu):
    from opymize_amd.linear import GradientOp, normest
    from oracle import pdhg
    op = GradientOp((20, 24), 2, scheme="forward")
    np.random.seed(11)
    e_gpu, it_gpu = normest(op)
    np.random.seed(11)
    e_cpu, it_cpu = pdhg.normest(ops.grad_matrix((20, 24), 2))
    assert abs(e_gpu - e_cpu) < 1e-9 * e_cpu and it_gpu == it_cpu
    # forward gradient on a large 2-D grid: |A|^2 -> 8
    assert 7.0 < e_gpu ** 2 <= 8.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_copy_sizes_and_alignments(gpu, dtype):
    """opy_copy: 16-byte chunks plus a tail when both arrays are 16-byte
    aligned, one element per thread otherwise; every size and offset gives
    the source bit for bit and writes nothing past the end."""
    import torch
    from opymize_amd import _hip
    rng = np.random.default_rng(12)
    src = torch.as_tensor(rng.standard_normal(2 ** 20 + 64).astype(dtype)).cuda()
    for n in (0, 1, 3, 4, 5, 255, 1023, 1025, 2 ** 20 + 3):
        for xo, yo in ((0, 0), (1, 1), (0, 1), (2, 0)):
            dst = torch.full((n + 16,), -7.0, dtype=src.dtype, device="cuda")
            _hip.call("copy", src.dtype, _hip.addr(src, xo), _hip.addr(dst, yo), n, _hip.stream())
            d = dst.cpu().numpy()
            assert np.array_equal(d[yo:yo + n], src[xo:xo + n].cpu().numpy()), (n, xo, yo)
            assert (d[:yo] == -7).all() and (d[yo + n:] == -7).all(), (n, xo, yo)


def test_staged_host_arrays(gpu):
    """numpy inputs are staged through the GPU kernels (same results)."""
    from opymize_amd.linear import GradientOp
    op = GradientOp((6, 7), 2, scheme="forward")
    x = np.random.default_rng(9).standard_normal(op.x.size)
    y = op.y.new()
    op(x, y)
    np.testing.assert_allclose(y, ops.grad_matrix((6, 7), 2).dot(x), atol=1e-14)


def _lifting_ops(N, seed=5):
    """The operators and shapes of the reference's tests/einsum.py:8-20
    (L = 50 labels, M = 60 regions, S = 2), with N points."""
    from opymize_amd.linear import IndexedMultAdj, MatrixMultRBatched, TangledMatrixMultR
    rng = np.random.default_rng(seed)
    L, M, S = 50, 60, 2
    A = rng.standard_normal((M, S, S))
    B = rng.standard_normal((M, S, S + 1))
    P = rng.integers(0, L, size=(M, S + 1))
    P[3, 1] = P[3, 0]  # a repeated label in one region
    tmp = rng.standard_normal((M, S + 1, 1, 1))
    return rng, [
        (IndexedMultAdj(L, N, P, B), ops.indexedmult_matrix(B, P, L, N).T),
        (MatrixMultRBatched(N, A), ops.matmul_r_batched_matrix(A, N)),
        (TangledMatrixMultR(N, tmp), ops.tangled_matmul_r_matrix(tmp, N)),
    ]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("N", [64, 4096])
def test_lifting_operators(gpu, dtype, N):
    """IndexedMultAdj, MatrixMultRBatched and TangledMatrixMultR and their
    adjoints against the oracle's entry-by-entry matrices, at the reference
    test's shapes (N = 64*64 there): y (+)= A x, the adjoint, rowwise l^1."""
    rng, cases = _lifting_ops(N)
    for op, Mtx in cases:
        for o, Mo in ((op, Mtx), (op.adjoint, Mtx.T.tocsr())):
            assert (o.y.size, o.x.size) == Mo.shape
            x = rng.standard_normal(o.x.size)
            y0 = rng.standard_normal(o.y.size)
            yd = dev(y0, dtype)
            o(dev(x, dtype), yd, add=True)
            close(host(yd), y0.astype(dtype) + Mo.dot(x.astype(dtype).astype(np.float64)), dtype, scale=10)
            yd = dev(y0, dtype)
            o(dev(x, dtype), yd)
            close(host(yd), Mo.dot(x.astype(dtype).astype(np.float64)), dtype, scale=10)
            if N == 64:
                r = dev(np.zeros(o.y.size), dtype)
                o.rowwise_lp(r)
                close(host(r), ops.rowwise_lp(Mo, 1), dtype, scale=10)
                r = dev(np.ones(o.y.size), dtype)
                o.rowwise_lp(r, p=2, add=True)
                close(host(r), 1.0 + ops.rowwise_lp(Mo, 2), dtype, scale=10)


def test_indexedmult_refuses_bad_indices(gpu):
    from opymize_amd.linear import IndexedMult
    with pytest.raises(ValueError, match="indices"):
        IndexedMult(4, 10, np.ones((2, 1, 3)), np.array([[0, 1, 4], [0, 0, 0]]))
    with pytest.raises(ValueError):
        IndexedMult(4, 10, np.ones((2, 1, 3)), np.zeros((2, 2), dtype=int))
