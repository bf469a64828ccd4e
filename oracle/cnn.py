"""FP64 reference of the warm-start network's 3x3 convolutions, forward and
backward, with the operands rounded to BF16 exactly where the HIP kernels
round them (activations and weights entering a matrix-core product).
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by: a direct per-tap loop against scipy.signal.correlate, and
finite-difference checks of the backward pass (tests/test_oracle_cnn.py)."""
import numpy as np


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(x, dy, dx):
    """x[n, h+dy-1, w+dx-1, c] with zero padding (NHWC)."""
    N, H, W, C = x.shape
    out = np.zeros_like(x)
    hs, ws = slice(max(0, 1 - dy), min(H, H + 1 - dy)), slice(max(0, 1 - dx), min(W, W + 1 - dx))
    ht, wt = slice(max(0, dy - 1), min(H, H + dy - 1)), slice(max(0, dx - 1), min(W, W + dx - 1))
    out[:, hs, ws, :] = x[:, ht, wt, :]
    return out


def conv_fwd(x, Wm, b=None, relu=False, round_ops=True):
    """y[n,h,w,co] = b[co] + sum_{tap,ci} W[tap,ci,co] x[n,h+dy-1,w+dx-1,ci]; W master layout (9, Ci, Co)."""
    xr = bf16(x) if round_ops else x.astype(np.float64)
    Wr = bf16(Wm) if round_ops else Wm.astype(np.float64)
    y = np.zeros(x.shape[:3] + (Wm.shape[2],))
    for tap in range(9):
        y += np.einsum("nhwc,cd->nhwd", shift(xr, tap // 3, tap % 3), Wr[tap])
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv_bwd_data(g, Wm, round_ops=True):
    """dL/dx for y = conv(x): sum_{tap,co} W[tap,ci,co] g[n,h-dy+1,w-dx+1,co]."""
    gr = bf16(g) if round_ops else g.astype(np.float64)
    Wr = bf16(Wm) if round_ops else Wm.astype(np.float64)
    dx = np.zeros(g.shape[:3] + (Wm.shape[1],))
    for tap in range(9):
        dy_, dx_ = tap // 3, tap % 3
        dx += np.einsum("nhwd,cd->nhwc", shift(gr, 2 - dy_, 2 - dx_), Wr[tap])
    return dx


def conv_bwd_weights(x, g, round_ops=True):
    """dW[tap,ci,co] = sum x[n,h+dy-1,w+dx-1,ci] g[n,h,w,co]; db[co] = sum g (FP32 g, not rounded)."""
    xr = bf16(x) if round_ops else x.astype(np.float64)
    gr = bf16(g) if round_ops else g.astype(np.float64)
    dW = np.stack([np.einsum("nhwc,nhwd->cd", shift(xr, t // 3, t % 3), gr) for t in range(9)])
    return dW, g.astype(np.float64).sum(axis=(0, 1, 2))


def net_forward(x, params, round_ops=True):
    """params: list of (W (9,32,32), b (32,)); ReLU after every layer but the last.  Returns activations."""
    acts = [x]
    for l, (Wm, b) in enumerate(params):
        acts.append(conv_fwd(acts[-1], Wm, b, relu=l < len(params) - 1, round_ops=round_ops))
    return acts


def net_loss_grad(x, tgt, params, wts=(1.0, 1.0, 1.0), nc=3, round_ops=True):
    """loss = sum_{p, c<nc} w_c (out - tgt)^2 / npix and its gradients (FP64)."""
    acts = net_forward(x, params, round_ops)
    out = acts[-1]
    npix = out.shape[0] * out.shape[1] * out.shape[2]
    w = np.zeros(out.shape[-1])
    w[:nc] = wts[:nc]
    d = out - tgt
    loss = float((w * d ** 2).sum() / npix)
    g = 2 * w * d / npix
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        if l < len(params) - 1:
            g = g * (acts[l + 1] > 0)
        dW, db = conv_bwd_weights(acts[l], g, round_ops)
        grads[l] = (dW, db)
        g = conv_bwd_data(g, params[l][0], round_ops)
    return loss, grads
