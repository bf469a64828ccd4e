"""Warm-start network kernels against the CPU oracle (BF16 operands rounded
identically, FP32 accumulation on the GPU vs FP64 in the oracle).
Tolerance: 2e-3 relative to the output scale (FP32 accumulation order over
K = 288 products plus BF16 re-rounding of intermediate activations)."""
import numpy as np
import pytest
import torch

from oracle import cnn

pytestmark = pytest.mark.gpu


def rel_close(a, b, tol=2e-3):
    scale = max(np.abs(b).max(), 1e-6)
    err = np.abs(a - b).max() / scale
    assert err < tol, "relative error %g" % err


@pytest.mark.parametrize("shape", [(1, 16, 16), (2, 21, 37), (3, 5, 70)])
def test_conv_forward_and_bias_relu(gpu, shape):
    from opymize_amd.warmstart.cnn import ConvNet, _call
    from opymize_amd import _hip
    N, H, W = shape
    rng = np.random.default_rng(0)
    net = ConvNet(layers=1, n_in=32, n_out=32, seed=3)
    x = rng.standard_normal((N, H, W, 32)).astype(np.float32)
    Wm = net.W(0).cpu().numpy().reshape(9, 32, 32)
    b = rng.standard_normal(32).astype(np.float32)
    net.b(0).copy_(torch.as_tensor(b))
    net._dirty = True
    net._prep()
    xd = torch.as_tensor(x).cuda()
    for act in (0, 1):
        y = torch.empty_like(xd)
        _call("opy_conv3x3_fwd_bf16", _hip.ptr(xd), _hip.ptr(y), _hip.ptr(net._wf[0]), _hip.ptr(net.b(0)), N, H, W,
              act, net.grid, _hip.stream())
        ref = cnn.conv_fwd(x, Wm, b, relu=bool(act))
        rel_close(y.cpu().numpy(), ref)


@pytest.mark.parametrize("shape", [(1, 16, 16), (2, 19, 33)])
def test_conv_backward_data_and_weights(gpu, shape):
    from opymize_amd.warmstart.cnn import ConvNet, _call
    from opymize_amd import _hip
    N, H, W = shape
    rng = np.random.default_rng(1)
    net = ConvNet(layers=1, n_in=32, n_out=32, seed=4)
    net._prep()
    Wm = net.W(0).cpu().numpy().reshape(9, 32, 32)
    x = rng.standard_normal((N, H, W, 32)).astype(np.float32)
    g = rng.standard_normal((N, H, W, 32)).astype(np.float32)
    gd, xd = torch.as_tensor(g).cuda(), torch.as_tensor(x).cuda()
    dx = torch.empty_like(gd)
    _call("opy_conv3x3_fwd_bf16", _hip.ptr(gd), _hip.ptr(dx), _hip.ptr(net._wb[0]), None, N, H, W, 0, net.grid,
          _hip.stream())
    rel_close(dx.cpu().numpy(), cnn.conv_bwd_data(g, Wm))
    nb = 7
    part = torch.empty(nb * net.per, dtype=torch.float32, device="cuda")
    _call("opy_conv3x3_wgrad_bf16", _hip.ptr(xd), _hip.ptr(gd), _hip.ptr(part), N, H, W, nb, _hip.stream())
    tot = torch.empty(net.per, dtype=torch.float32, device="cuda")
    _call("opy_sum_partials_f32", _hip.ptr(part), nb, net.per, _hip.ptr(tot), _hip.stream())
    t = tot.cpu().numpy()
    dW_ref, db_ref = cnn.conv_bwd_weights(x, g)
    rel_close(t[:9 * 1024].reshape(9, 32, 32).transpose(0, 2, 1), dW_ref)
    rel_close(t[9 * 1024:], db_ref, tol=1e-5)


def test_network_loss_and_gradients_match_oracle(gpu):
    from opymize_amd.warmstart.cnn import ConvNet
    rng = np.random.default_rng(2)
    N, H, W = 2, 18, 20
    net = ConvNet(layers=3, n_in=2, n_out=3, seed=5)
    x = np.zeros((N, H, W, 32), dtype=np.float32)
    x[..., :2] = rng.standard_normal((N, H, W, 2))
    tgt = np.zeros_like(x)
    tgt[..., :3] = rng.standard_normal((N, H, W, 3))
    params = []
    for l in range(3):
        params.append((net.W(l).cpu().numpy().reshape(9, 32, 32).astype(np.float64),
                       net.b(l).cpu().numpy().astype(np.float64)))
    xd, td = torch.as_tensor(x).cuda(), torch.as_tensor(tgt).cuda()
    out, acts = net.forward(xd, keep=True)
    part, g = net.loss_grad(out, td, (1.0, 0.5, 0.5))
    net.backward(acts, g)
    loss_ref, grads_ref = cnn.net_loss_grad(x, tgt, params, wts=(1.0, 0.5, 0.5), nc=3)
    assert abs(part.sum().item() - loss_ref) < 2e-3 * abs(loss_ref)
    G = net.grads.cpu().numpy()
    for l in range(3):
        gW = G[l * net.per:l * net.per + 9 * 1024].reshape(9, 32, 32)
        gb = G[l * net.per + 9 * 1024:(l + 1) * net.per]
        rel_close(gW, grads_ref[l][0], tol=1e-2)
        rel_close(gb, grads_ref[l][1], tol=1e-2)


def test_training_reduces_loss(gpu):
    """Fit a fixed smoothing target: loss must drop by 10x within 300 Adam steps."""
    from opymize_amd.warmstart.cnn import ConvNet
    rng = np.random.default_rng(3)
    N, H, W = 8, 32, 32
    net = ConvNet(layers=3, n_in=2, n_out=3, seed=6)
    x = np.zeros((N, H, W, 32), dtype=np.float32)
    x[..., 0] = rng.standard_normal((N, H, W))
    tgt = np.zeros_like(x)
    k = np.ones((3, 3)) / 9.0
    from scipy.signal import correlate
    for n in range(N):
        xn = x[n, ..., 0].astype(np.float64)
        tgt[n, ..., 0] = correlate(xn, k, mode="same") - xn
    xd, td = torch.as_tensor(x).cuda(), torch.as_tensor(tgt).cuda()
    losses = [net.train_step(xd, td, lr=3e-3).sum().item() for _ in range(300)]
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


def test_trained_network_gain_over_cold_start(gpu):
    """What the committed network (trained on solver outputs of the rectangle
    generator, seed 11) gains on held-out instances (seed 12): a start point
    with a finite, small relative gap (cold: y = 0 gives a dual value near 0),
    and fewer iterations to a relative gap of 1e-2.  Both runs then reach the
    same solution: ||x_warm - x_cold|| <= sqrt(2 g_w / lam) + sqrt(2 g_c / lam)
    (strong convexity of the SSD term)."""
    import os
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opymize_amd", "warmstart",
                        "weights_tv_l2.npz")
    specs = synth.batch_square(8, side=256, seed=12)
    imgs = synth.images(specs)
    lams = [s["lam"] for s in specs]
    ws = WarmStart(ConvNet.load(path))
    out = {}
    for mode in ("cold", "warm"):
        bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, 8))
        if mode == "warm":
            ws.predict_into(bat, lams)
        r0 = bat.solve(steps="adaptive", step_bound=0.124, tol=1e-30, max_iters=0, check_start=True)
        first = np.full(len(imgs), np.inf)

        def cb(b, it):
            rg = np.array([abs(q["relgap"]) for q in b.results()])
            first[(rg < 1e-2) & ~np.isfinite(first)] = it
            return bool(np.all(np.isfinite(first)) and it >= 2000)

        bat.solve(steps="adaptive", step_bound=0.124, tol=1e-30, max_iters=2000, block_iters=50, callback=cb)
        res = bat.results()
        out[mode] = (np.array([abs(q["relgap"]) for q in r0]), first, bat.solution()[0],
                     np.array([q["gap"] for q in res]))
    start_c, first_c, x_c, gap_c = out["cold"]
    start_w, first_w, x_w, gap_w = out["warm"]
    assert np.all(start_w < 10.0) and np.all(start_c > 1e6)
    assert np.all(np.isfinite(first_w)) and np.all(np.isfinite(first_c))
    assert np.median(first_w) < np.median(first_c)
    assert first_w.sum() < first_c.sum()
    for b, lam in enumerate(lams):
        bound = np.sqrt(2 * max(gap_w[b], 0) / lam) + np.sqrt(2 * max(gap_c[b], 0) / lam)
        assert np.linalg.norm(x_w[b] - x_c[b]) <= bound * 1.001 + 1e-3


@pytest.mark.parametrize("shape", [(2, 21, 37), (1, 64, 48)])
def test_bf16_activations_give_the_same_output(gpu, shape):
    """Inference with BF16 activations between layers returns bitwise the
    output of the float32-activation forward pass: every layer rounds its
    operands to BF16 for the matrix cores, so storing them as BF16 changes no
    operand.  Also the float32 -> float32 mode of opy_conv3x3_fwd_io equals
    opy_conv3x3_fwd_bf16."""
    import os
    from opymize_amd.warmstart.cnn import ConvNet, _call
    from opymize_amd import _hip
    N, H, W = shape
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    net = ConvNet.load(os.path.join(root, "opymize_amd", "warmstart", "weights_tv_l2.npz"))
    rng = np.random.default_rng(4)
    x = np.zeros((N, H, W, 32), dtype=np.float32)
    x[..., 0] = rng.random((N, H, W)) - 0.5
    x[..., 1] = 0.7
    xd = torch.as_tensor(x).cuda()
    ref = net.forward(xd).cpu().numpy()
    out = net.forward(xd, bf16_acts=True).cpu().numpy()
    np.testing.assert_array_equal(out, ref)
    y1, y2 = torch.empty_like(xd), torch.empty_like(xd)
    _call("opy_conv3x3_fwd_bf16", _hip.ptr(xd), _hip.ptr(y1), _hip.ptr(net._wf[0]), _hip.ptr(net.b(0)), N, H, W, 1,
          net.grid, _hip.stream())
    _call("opy_conv3x3_fwd_io", _hip.ptr(xd), _hip.ptr(y2), _hip.ptr(net._wf[0]), _hip.ptr(net.b(0)), N, H, W, 1,
          0, 0, net.grid, _hip.stream())
    np.testing.assert_array_equal(y1.cpu().numpy(), y2.cpu().numpy())


@pytest.mark.parametrize("sizes", [[(40, 56)] * 5, [(17, 33), (40, 56), (17, 33), (64, 64), (17, 33)]])
def test_fused_warm_start_io_equals_encoded_path(gpu, sizes):
    """The warm start through opy_conv3x3_first / opy_conv3x3_last (input
    encoding and output decoding inside the first and last layers) writes
    bitwise the same starts as the path through the encoded 32-channel
    tensors, for a chunk of equally spaced instances and for an uneven batch
    (whose same-shape instances are not equally spaced)."""
    import os
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opymize_amd", "warmstart",
                        "weights_tv_l2.npz")
    imgs = [synth.rects(H, W, seed=50 + i) for i, (H, W) in enumerate(sizes)]
    lams = [1.0, 2.5, 7.0, 13.0, 19.5][:len(imgs)]
    ws = WarmStart(ConvNet.load(path), chunk=3)
    out = {}
    for fused in (False, True):
        bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, 6))
        ws.predict_into(bat, lams, fused=fused)
        out[fused] = (bat.x[0].cpu().numpy(), bat.y[0].cpu().numpy())
    np.testing.assert_array_equal(out[True][0], out[False][0])
    np.testing.assert_array_equal(out[True][1], out[False][1])
    assert np.abs(out[True][0]).max() > 0.1


@pytest.mark.parametrize("fused", [False, True])
def test_warm_start_dual_lies_in_each_instance_ball(gpu, fused):
    """The predicted dual start is projected onto |y| <= tvw of each instance
    (the batch's own TV weights, here unlike the network's 1), so every
    instance starts dual-feasible."""
    import os
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opymize_amd", "warmstart",
                        "weights_tv_l2.npz")
    imgs = [synth.rects(40, 56, seed=70 + i) for i in range(4)]
    lams = [1.5, 4.0, 9.0, 16.0]
    tvw = [0.05, 0.05, 2.0, 0.3]
    bat = TVBatch(imgs, lams, "l2", tvw=tvw, precision="float", tile=(64, 6))
    WarmStart(ConvNet.load(path), chunk=4).predict_into(bat, lams, fused=fused)
    _, ys = bat.solution()
    for b in range(4):
        n = np.linalg.norm(ys[b].reshape(-1, 2), axis=1)
        assert n.max() <= tvw[b] * (1 + 1e-6), (b, n.max(), tvw[b])
    # the small radius is active: the network's output exceeds it somewhere
    assert np.linalg.norm(ys[0].reshape(-1, 2), axis=1).max() > 0.9 * tvw[0]


def test_warm_start_at_full_size_matches_oracle_windows(gpu):
    """The bench's inference path at its real shape: the packed warm start of
    two 1024x1024 instances (first and last layers reading and writing the
    packed buffers, BF16 activations) against the oracle network on windows
    of the same images.  Six 3x3 layers see 6 pixels around an output pixel,
    so the oracle runs on a window grown by 6 pixels (clipped at the image
    edge, where both sides pad with zeros) and its interior is compared.
    Windows at a corner, an edge and the middle."""
    import os
    from opymize_amd.solvers.tv import TVBatch
    from opymize_amd.tools import synth
    from opymize_amd.warmstart.cnn import ConvNet, WarmStart
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opymize_amd", "warmstart",
                        "weights_tv_l2.npz")
    net = ConvNet.load(path)
    imgs = [synth.rects(1024, 1024, seed=80 + i) for i in range(2)]
    lams = [2.0, 9.0]
    bat = TVBatch(imgs, lams, "l2", precision="float", tile=(64, 6))
    WarmStart(net, chunk=2).predict_into(bat, lams)
    xs, ys = bat.solution()
    params = [(net.W(l).cpu().numpy().reshape(9, 32, 32).astype(np.float64),
               net.b(l).cpu().numpy().astype(np.float64)) for l in range(net.layers)]
    m = net.layers
    for b in range(2):
        f = imgs[b].astype(np.float32)
        X = xs[b].reshape(1024, 1024)
        Y = ys[b].reshape(1024, 1024, 2)
        for r0, c0 in ((0, 0), (0, 500), (480, 600), (1024 - 40, 1024 - 48)):
            h, w = 40, 48
            a0, a1 = max(0, r0 - m), min(1024, r0 + h + m)
            b0, b1 = max(0, c0 - m), min(1024, c0 + w + m)
            x = np.zeros((1, a1 - a0, b1 - b0, 32), dtype=np.float32)
            x[0, ..., 0] = f[a0:a1, b0:b1] - np.float32(0.5)
            x[0, ..., 1] = np.float32(np.log(np.float32(lams[b])) / 3.0)
            out = cnn.net_forward(x, params)[-1][0]
            win = out[r0 - a0:r0 - a0 + h, c0 - b0:c0 - b0 + w]
            x_ref = f[r0:r0 + h, c0:c0 + w].astype(np.float64) + win[..., 0]
            y_ref = win[..., 1:3]
            nrm = np.linalg.norm(y_ref, axis=-1, keepdims=True)
            y_ref = y_ref * np.minimum(1.0, 1.0 / np.maximum(nrm, 1e-300))  # |y| <= tvw = 1
            rel_close(X[r0:r0 + h, c0:c0 + w], x_ref)
            rel_close(Y[r0:r0 + h, c0:c0 + w], y_ref)
