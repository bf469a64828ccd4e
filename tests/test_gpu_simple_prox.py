"""GPU parity of the simple prox family and the affine / indicator
functionals (reference operators/proj.py:19-85, operators/affine.py:15-146,
functionals/affine.py:9-90, functionals/indicators.py) against their closed
forms, and the operator test helpers (reference tools/tests.py:47-95) run on
every linear operator of the package.

Tolerances: FP64 to 1e-12 (these maps are single roundings: max, min, a copy,
one fused multiply-add), FP32 to 2e-6 relative; functional values to 1e-10
(FP64) and 1e-6 (FP32) relative, summed in FP64 on the device."""
import numpy as np
import pytest
import scipy.sparse as sp

from opymize_amd.tools import tests as optests

pytestmark = pytest.mark.gpu

DT = {"double": np.float64, "float": np.float32}


def dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


def host(t):
    return t.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("type_t", ["double", "float"])
def test_pos_neg_proj(gpu, type_t):
    from opymize_amd.operators import PosProj, NegProj
    N = 10007
    # T(x) = max(0, x) / min(0, x) (reference proj.py:19-85)
    optests.test_gpu_op(PosProj(N), lambda x, y0: np.maximum(x, 0.0), type_t, seed=1)
    optests.test_gpu_op(NegProj(N), lambda x, y0: np.minimum(x, 0.0), type_t, seed=2)
    # in place (y is None), as the solver calls prox maps
    x = np.random.default_rng(3).standard_normal(N).astype(DT[type_t])
    xd = dev(x, DT[type_t])
    PosProj(N)(xd)
    assert np.array_equal(host(xd), np.maximum(x, 0).astype(np.float64))
    xd = dev(x, DT[type_t])
    NegProj(N)(xd)
    assert np.array_equal(host(xd), np.minimum(x, 0).astype(np.float64))


@pytest.mark.parametrize("type_t", ["double", "float"])
def test_const_and_constrain(gpu, type_t):
    from opymize_amd.operators import ConstOp, ConstrainOp
    rng = np.random.default_rng(4)
    N, M = 999, 3
    c = rng.standard_normal((N, M))
    cc = c.astype(DT[type_t]).astype(np.float64)
    # ConstOp: y = c whatever x is (reference affine.py:15-46); x has its own size
    op = ConstOp(17, c)
    optests.test_gpu_op(op, lambda x, y0: cc.ravel(), type_t, seed=5)
    # ConstrainOp: rows in the mask take c, the others keep x (reference affine.py:48-83)
    mask = rng.random(N) < 0.4
    op = ConstrainOp(mask, c)

    def ref(x, y0):
        out = x.reshape(N, M).copy()
        out[mask, :] = cc[mask, :]
        return out

    optests.test_gpu_op(op, ref, type_t, seed=6)
    # in place
    x = rng.standard_normal(N * M)
    xd = dev(x, DT[type_t])
    op(xd)
    np.testing.assert_array_equal(host(xd), ref(x.astype(DT[type_t]).astype(np.float64), None).ravel())


@pytest.mark.parametrize("type_t", ["double", "float"])
@pytest.mark.parametrize("kind", ["array_shift", "scalar_shift", "array_ab", "scalar_all"])
def test_shiftscale(gpu, type_t, kind):
    """T(x) = a*(x + b*shift) with every combination of arrays and scalars
    (reference affine.py:85-146), with a non-zero shift in each case."""
    from opymize_amd.operators import ShiftScaleOp
    rng = np.random.default_rng(7)
    N = 4099
    arr = lambda: rng.uniform(-2, 2, N)
    shift, a, b = {
        "array_shift": (arr(), 0.7, -1.3),
        "scalar_shift": (0.45, arr(), 2.0),
        "array_ab": (arr(), arr(), arr()),
        "scalar_all": (-0.8, 1.6, 0.3),
    }[kind]
    r = lambda v: v.astype(DT[type_t]).astype(np.float64) if isinstance(v, np.ndarray) else float(
        np.asarray(v, dtype=DT[type_t]))
    op = ShiftScaleOp(N, shift, a, b)
    tol = 1e-12 if type_t == "double" else 4e-6
    optests.test_gpu_op(op, lambda x, y0: r(a) * (x + r(b) * r(shift)), type_t, rtol=tol, atol=tol, seed=8)


def test_shiftscale_and_const_refuse_wrong_sizes(gpu):
    from opymize_amd.operators import ShiftScaleOp, ConstOp, ConstrainOp
    N = 100
    with pytest.raises(ValueError, match="shift has 99 entries"):
        ShiftScaleOp(N, np.ones(99), 1.0, 1.0)(dev(np.zeros(N), np.float64), dev(np.zeros(N), np.float64))
    with pytest.raises(ValueError, match="ConstOp"):
        ConstOp(N, np.ones(50))(dev(np.zeros(N), np.float64), dev(np.zeros(60), np.float64))
    with pytest.raises(ValueError, match="ConstrainOp"):
        ConstrainOp(np.ones(10, bool), np.ones((10, 2)))(dev(np.zeros(21), np.float64))


@pytest.mark.parametrize("type_t", ["double", "float"])
def test_affine_and_indicator_values(gpu, type_t):
    """Values and infeasibilities (reference functionals/affine.py:9-90,
    indicators.py:7-57) and the prox maps of the affine functionals."""
    from opymize_amd.functionals import AffineFct, MaskedAffineFct, IndicatorFct, ConstrainFct, ZeroFct
    dt = DT[type_t]
    rt = 1e-10 if dt == np.float64 else 1e-6
    rng = np.random.default_rng(9)
    N, M = 800, 3
    x = rng.standard_normal(N * M).astype(dt).astype(np.float64)
    c = rng.standard_normal(N * M)
    cr = c.astype(dt).astype(np.float64)
    # AffineFct: b + <c, x>, array and scalar c
    v, i = AffineFct(N * M, b=2.5, c=c)(dev(x, dt))
    assert np.isclose(v, 2.5 + x.dot(cr), rtol=rt) and i == 0
    v, i = AffineFct(N * M, b=-1.0, c=0.75)(dev(x, dt))
    assert np.isclose(v, -1.0 + 0.75 * x.sum(), rtol=rt) and i == 0
    # its prox: x - tau c
    xd = dev(x, dt)
    AffineFct(N * M, b=2.5, c=c).prox(0.3)(xd)
    np.testing.assert_allclose(host(xd), x - 0.3 * cr, rtol=0, atol=4 * rt)
    # MaskedAffineFct: sum over masked rows of c*x, infeas = max |x| on the others
    mask = rng.random(N) < 0.6
    C2 = c.reshape(N, M)
    X2 = x.reshape(N, M)
    F = MaskedAffineFct(mask, C2)
    v, i = F(dev(x, dt))
    assert np.isclose(v, (X2[mask] * C2.astype(dt).astype(np.float64)[mask]).sum(), rtol=rt)
    assert np.isclose(i, np.abs(X2[~mask]).max(), rtol=rt)
    # its prox: scale*(x - tau c) with scale 1 on masked rows, 0 elsewhere
    xd = dev(x, dt)
    F.prox(0.4)(xd)
    want = np.where(mask[:, None], X2 - 0.4 * C2.astype(dt).astype(np.float64), 0.0).ravel()
    np.testing.assert_allclose(host(xd), want, rtol=0, atol=4 * rt)
    # the conjugate pair: ConstrainFct(mask, c) value 0, infeas max |x - c| on the masked rows
    v, i = F.conj(dev(x, dt))
    assert v == 0 and np.isclose(i, np.abs(X2 - C2.astype(dt).astype(np.float64))[mask].max(), rtol=rt)
    v, i = ConstrainFct(mask, C2)(dev(x, dt))
    assert v == 0 and np.isclose(i, np.abs(X2 - C2.astype(dt).astype(np.float64))[mask].max(), rtol=rt)
    # IndicatorFct: c2 if x == c1, infeas = max |x - c1|; conj = AffineFct(b=-c2, c=c1)
    v, i = IndicatorFct(N * M, c1=0.5, c2=1.25)(dev(x, dt))
    assert v == 1.25 and np.isclose(i, np.abs(x - 0.5).max(), rtol=rt)
    v, i = IndicatorFct(N * M, c1=c, c2=0.0).conj(dev(x, dt))
    assert np.isclose(v, x.dot(cr), rtol=rt)
    v, i = ZeroFct(N * M)(dev(x, dt))
    assert v == 0 and i == 0


def _linops(rng):
    """Every linear operator of the package on a small shape, with the
    tolerance of its arithmetic."""
    from opymize_amd.linear import (GradientOp, DivergenceOp, LaplacianOp, MatrixMult, MatrixMultR,
                                    DiagMatrixMultR, IndexedMult, TangledMatrixMultR,
                                    MatrixMultRBatched, ScaleOp, IdentityOp, ZeroOp, BlockOp, CsrOp, MatrixOp)
    ops = [
        ("grad_fwd", GradientOp((9, 11), 2, scheme="forward", imageh=(0.7, 1.3), weights=(1.0, 0.5))),
        ("grad_cen", GradientOp((6, 7, 5), 1, scheme="centered")),
        ("div", DivergenceOp((9, 11), 2, scheme="forward")),
        ("lapl_N", LaplacianOp((8, 9), 2, boundary="neumann")),
        ("lapl_C", LaplacianOp((8, 9), 1, boundary="curvature")),
        ("lapl_S", LaplacianOp((8, 9), 1, boundary="second-order")),
        ("matmul", MatrixMult(37, rng.standard_normal((5, 7)))),
        ("matmul_r", MatrixMultR(37, rng.standard_normal((7, 5)))),
        ("matmul_r16", MatrixMultR(70, rng.standard_normal((16, 16)))),
        ("diag_r", DiagMatrixMultR(37, rng.standard_normal(6))),
        ("indexed", IndexedMult(5, 11, rng.standard_normal((3, 2, 2)), rng.integers(0, 5, (3, 2)))),
        ("tangled", TangledMatrixMultR(13, rng.standard_normal((2, 3, 4, 2)))),
        ("batched", MatrixMultRBatched(13, rng.standard_normal((3, 4, 2)))),
        ("scale", ScaleOp(50, rng.standard_normal(50))),
        ("ident", IdentityOp(50)),
        ("csr", CsrOp(sp.random(40, 30, density=0.2, random_state=3, format="csr"))),
        ("dense", MatrixOp(rng.standard_normal((12, 9)))),
    ]
    g = GradientOp((5, 6), 1)
    ops.append(("block", BlockOp([[g, ZeroOp(g.y.size, 8)], [MatrixOp(rng.standard_normal((3, 30))),
                                                              MatrixOp(rng.standard_normal((3, 8)))]])))
    return ops


@pytest.mark.parametrize("name", [n for n, _ in _linops(np.random.default_rng(0))])
def test_helpers_on_every_linop(gpu, name):
    """The reference's own operator checks (tools/tests.py:47-67): the adjoint
    identity and the row-wise l1 norms against explicit rows A^T e_i, in FP64."""
    np.random.seed(10)
    op = dict(_linops(np.random.default_rng(0)))[name]
    optests.test_adjoint(op)
    optests.test_adjoint(op.adjoint)
    optests.test_rowwise_lp(op, p=1, maxiter=10)
    optests.test_rowwise_lp(op.adjoint, p=1, maxiter=10)
