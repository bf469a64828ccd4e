"""Pins of the CNN oracle (CPU only)."""
import numpy as np
from scipy.signal import correlate

from oracle import cnn


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.140625, 1e-3], dtype=np.float32)
    r = cnn.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.015625
    assert r[4] == -3.140625
    assert abs(r[5] - 1e-3) < 1e-3 * 2 ** -8


def test_conv_fwd_against_scipy_correlate():
    rng = np.random.default_rng(0)
    N, H, W, Ci, Co = 2, 7, 9, 3, 4
    x = rng.standard_normal((N, H, W, Ci))
    Wm = rng.standard_normal((9, Ci, Co))
    b = rng.standard_normal(Co)
    y = cnn.conv_fwd(x, Wm, b, round_ops=False)
    for n in range(N):
        for co in range(Co):
            ref = b[co] + sum(correlate(x[n, :, :, ci], Wm[:, ci, co].reshape(3, 3), mode="same")
                              for ci in range(Ci))
            np.testing.assert_allclose(y[n, :, :, co], ref, atol=1e-12)


def test_network_gradients_finite_differences():
    rng = np.random.default_rng(1)
    N, H, W, C = 1, 5, 6, 4
    params = [(0.4 * rng.standard_normal((9, C, C)), 0.1 * rng.standard_normal(C)) for _ in range(3)]
    x = rng.standard_normal((N, H, W, C))
    tgt = rng.standard_normal((N, H, W, C))
    loss, grads = cnn.net_loss_grad(x, tgt, params, wts=(1.0, 0.5, 2.0), nc=3, round_ops=False)
    eps = 1e-6
    for l in range(3):
        for idx in [(0, 0, 0), (4, 1, 2), (8, 3, 1)]:
            Wp = [(w.copy(), b.copy()) for w, b in params]
            Wp[l][0][idx] += eps
            lp, _ = cnn.net_loss_grad(x, tgt, Wp, wts=(1.0, 0.5, 2.0), nc=3, round_ops=False)
            Wp[l][0][idx] -= 2 * eps
            lm, _ = cnn.net_loss_grad(x, tgt, Wp, wts=(1.0, 0.5, 2.0), nc=3, round_ops=False)
            assert abs((lp - lm) / (2 * eps) - grads[l][0][idx]) < 1e-6 * max(1, abs(grads[l][0][idx]))
        for c in range(C):
            Wp = [(w.copy(), b.copy()) for w, b in params]
            Wp[l][1][c] += eps
            lp, _ = cnn.net_loss_grad(x, tgt, Wp, wts=(1.0, 0.5, 2.0), nc=3, round_ops=False)
            Wp[l][1][c] -= 2 * eps
            lm, _ = cnn.net_loss_grad(x, tgt, Wp, wts=(1.0, 0.5, 2.0), nc=3, round_ops=False)
            assert abs((lp - lm) / (2 * eps) - grads[l][1][c]) < 1e-6 * max(1, abs(grads[l][1][c]))
