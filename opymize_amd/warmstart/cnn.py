"""Warm-start network for the batched TV solver.

A 6-layer 3x3 convolutional network (32 channels, ReLU) maps the noisy image
and its data weight lambda to a starting point (x0, y0) of PDHG:
  input channels:  f - 0.5, log(lambda)/3        (zero padded to 32)
  output channels: x0 - f, y0_vertical, y0_horizontal  (y0 projected onto |y| <= tvw)
All layers run on the hand-written matrix-core kernels of libopyhip.so
(BF16 operands, FP32 accumulation, FP32 master weights, Adam).  It is
trained on pairs (f, PDHG solution) that the solver itself produces; the
solver always continues from the prediction and certifies the result with
the FP64 gap, falling back to a cold start for an instance whose gap does
not improve.
"""
import ctypes

import numpy as np
import torch

from opymize_amd import _hip

CH = 32
_bound = False


def _bind():
    global _bound
    if _bound:
        return
    lib = _hip.load()
    P, I, L, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    sigs = {
        "opy_conv3x3_fwd_bf16": [P, P, P, P, I, I, I, I, I, P],
        "opy_conv3x3_fwd_io": [P, P, P, P, I, I, I, I, I, I, I, P],
        "opy_conv3x3_first": [P, L, P, P, P, P, I, I, I, I, P],
        "opy_conv3x3_last": [P, P, L, P, P, P, P, I, I, I, I, P],
        "opy_conv3x3_wgrad_bf16": [P, P, P, I, I, I, I, P],
        "opy_sum_partials_f32": [P, I, I, P, P],
        "opy_conv3x3_prep_weights": [P, P, I, P],
        "opy_relu_bwd_f32": [P, P, L, P],
        "opy_adam_f32": [P, P, P, P, I, D, D, D, D, I, P],
        "opy_mse3_f32": [P, P, P, L, I, D, D, D, P, I, P],
    }
    for k, a in sigs.items():
        fn = getattr(lib, k)
        fn.restype, fn.argtypes = ctypes.c_int, a
    _bound = True


def _call(name, *args):
    lib = _hip.load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise _hip.HipError("%s failed: %s" % (name, lib.opy_last_error().decode()))


class ConvNet(object):
    """Stack of 3x3 conv layers with 32 -> 32 channels (ReLU between layers)."""

    def __init__(self, layers=6, n_in=2, n_out=3, seed=0):
        _hip.require_device()
        _bind()
        self.layers = layers
        rng = np.random.default_rng(seed)
        per = 9 * CH * CH + CH
        flat = np.zeros(layers * per, dtype=np.float32)
        for l in range(layers):
            fan_in = 9 * (n_in if l == 0 else CH)
            Wm = rng.standard_normal((9, CH, CH)) * np.sqrt(2.0 / fan_in)
            if l == 0:
                Wm[:, n_in:, :] = 0.0
            if l == layers - 1:
                Wm[:, :, n_out:] = 0.0
                Wm *= 0.1  # start close to "x0 = f, y0 = 0"
            flat[l * per:l * per + 9 * CH * CH] = Wm.astype(np.float32).ravel()
        self.per = per
        self.params = torch.as_tensor(flat).cuda()
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.n_in, self.n_out = n_in, n_out
        self._wf = [torch.empty(9 * CH * CH, dtype=torch.bfloat16, device="cuda") for _ in range(layers)]
        self._wb = [torch.empty(9 * CH * CH, dtype=torch.bfloat16, device="cuda") for _ in range(layers)]
        self.grid = 8 * torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        self._dirty = True
        self._ibufs = {}  # inference buffers by input shape: two BF16 activations and the float32 output

    # -- parameters -------------------------------------------------------
    def W(self, l):
        return self.params[l * self.per:l * self.per + 9 * CH * CH]

    def b(self, l):
        return self.params[l * self.per + 9 * CH * CH:(l + 1) * self.per]

    def _prep(self):
        if not self._dirty:
            return
        s = _hip.stream()
        for l in range(self.layers):
            _call("opy_conv3x3_prep_weights", _hip.ptr(self.W(l)), _hip.ptr(self._wf[l]), 0, s)
            _call("opy_conv3x3_prep_weights", _hip.ptr(self.W(l)), _hip.ptr(self._wb[l]), 1, s)
        self._dirty = False

    def state_dict(self):
        return {"params": self.params.cpu().numpy(), "layers": self.layers, "n_in": self.n_in, "n_out": self.n_out}

    def load_state_dict(self, d):
        self.params.copy_(torch.as_tensor(np.asarray(d["params"], dtype=np.float32)))
        self._dirty = True

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path):
        d = np.load(path)
        net = cls(layers=int(d["layers"]), n_in=int(d["n_in"]), n_out=int(d["n_out"]))
        net.load_state_dict(d)
        return net

    # -- passes ----------------------------------------------------------------
    def forward(self, x, keep=False, bufs=None, bf16_acts=False):
        """x: (N, H, W, 32) float32 device tensor.  Returns the output (and the
        list of activations when keep=True, needed by backward).

        bf16_acts=True (inference): the activations between layers are kept in
        BF16 (the layers round their operands to BF16 in any case, so the
        output is the same) in buffers owned by the network; the output is
        float32 and valid until the next such call."""
        self._prep()
        N, H, W, C = x.shape
        assert C == CH and x.dtype == torch.float32 and x.is_contiguous()
        s = _hip.stream()
        if bf16_acts and not keep:
            key = tuple(x.shape)
            if key not in self._ibufs:
                self._ibufs = {key: (torch.empty(key, dtype=torch.bfloat16, device="cuda"),
                                     torch.empty(key, dtype=torch.bfloat16, device="cuda"),
                                     torch.empty(key, dtype=torch.float32, device="cuda"))}
            b16a, b16b, outf = self._ibufs[key]
            cur, cur16 = x, 0
            for l in range(self.layers):
                last = l == self.layers - 1
                out = outf if last else (b16a, b16b)[l % 2]
                _call("opy_conv3x3_fwd_io", _hip.ptr(cur), _hip.ptr(out), _hip.ptr(self._wf[l]), _hip.ptr(self.b(l)),
                      N, H, W, 0 if last else 1, cur16, 0 if last else 1, self.grid, s)
                cur, cur16 = out, 1
            return outf
        acts = [x]
        cur = x
        for l in range(self.layers):
            if keep or bufs is None:
                out = torch.empty_like(x)
            else:
                out = bufs[l % 2]
            relu = 1 if l < self.layers - 1 else 0
            _call("opy_conv3x3_fwd_bf16", _hip.ptr(cur), _hip.ptr(out), _hip.ptr(self._wf[l]), _hip.ptr(self.b(l)),
                  N, H, W, relu, self.grid, s)
            if keep:
                acts.append(out)
            cur = out
        return (cur, acts) if keep else cur

    def forward_packed(self, f_ptr, stride, loglam, N, H, W, xo_ptr, yo_ptr):
        """Inference for N images of H x W at f_ptr + n*stride floats (the
        solver's packed FP32 images) straight into the packed FP32 starts
        x (xo_ptr + n*stride) and y (yo_ptr + 2n*stride): the same values as
        forward(encoded input, bf16_acts=True) followed by the decoding
        x = ((f - 0.5) + 0.5) + out0, y = (out1, out2), without the
        32-channel FP32 input and output tensors.  loglam: (N,) float32
        device tensor, the input's second channel."""
        self._prep()
        assert self.layers >= 2 and loglam.dtype == torch.float32 and loglam.numel() == N
        s = _hip.stream()
        key = (N, H, W, CH)
        if key not in self._ibufs:
            self._ibufs = {key: (torch.empty(key, dtype=torch.bfloat16, device="cuda"),
                                 torch.empty(key, dtype=torch.bfloat16, device="cuda"),
                                 torch.empty(key, dtype=torch.float32, device="cuda"))}
        b16 = self._ibufs[key][:2]
        _call("opy_conv3x3_first", f_ptr, int(stride), _hip.ptr(loglam), _hip.ptr(b16[0]), _hip.ptr(self._wf[0]),
              _hip.ptr(self.b(0)), N, H, W, self.grid, s)
        for l in range(1, self.layers - 1):
            _call("opy_conv3x3_fwd_io", _hip.ptr(b16[(l - 1) % 2]), _hip.ptr(b16[l % 2]), _hip.ptr(self._wf[l]),
                  _hip.ptr(self.b(l)), N, H, W, 1, 1, 1, self.grid, s)
        l = self.layers - 1
        _call("opy_conv3x3_last", _hip.ptr(b16[(l - 1) % 2]), f_ptr, int(stride), xo_ptr, yo_ptr,
              _hip.ptr(self._wf[l]), _hip.ptr(self.b(l)), N, H, W, self.grid, s)

    def backward(self, acts, gout):
        """Gradients of all parameters into self.grads given dL/d(output)."""
        N, H, W, _ = gout.shape
        s = _hip.stream()
        nb = min(self.grid, N * ((H + 15) // 16) * ((W + 15) // 16))
        part = torch.empty(nb * self.per, dtype=torch.float32, device="cuda")
        g = gout
        for l in range(self.layers - 1, -1, -1):
            if l < self.layers - 1:
                _call("opy_relu_bwd_f32", _hip.ptr(g), _hip.ptr(acts[l + 1]), g.numel(), s)
            _call("opy_conv3x3_wgrad_bf16", _hip.ptr(acts[l]), _hip.ptr(g), _hip.ptr(part), N, H, W, nb, s)
            gl = self.grads[l * self.per:(l + 1) * self.per]
            # partials hold dW as [tap][co][ci]; reorder to the master [tap][ci][co] below
            tmp = torch.empty(self.per, dtype=torch.float32, device="cuda")
            _call("opy_sum_partials_f32", _hip.ptr(part), nb, self.per, _hip.ptr(tmp), s)
            gl[:9 * CH * CH].copy_(tmp[:9 * CH * CH].view(9, CH, CH).transpose(1, 2).reshape(-1))
            gl[9 * CH * CH:].copy_(tmp[9 * CH * CH:])
            if l > 0:
                gn = torch.empty_like(g)
                _call("opy_conv3x3_fwd_bf16", _hip.ptr(g), _hip.ptr(gn), _hip.ptr(self._wb[l]), None, N, H, W, 0,
                      self.grid, s)
                g = gn
        return self.grads

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _call("opy_adam_f32", _hip.ptr(self.params), _hip.ptr(self.grads), _hip.ptr(self.m), _hip.ptr(self.v),
              self.params.numel(), lr, b1, b2, eps, self.step, _hip.stream())
        self._dirty = True

    def loss_grad(self, out, tgt, wts=(1.0, 1.0, 1.0)):
        N, H, W, _ = out.shape
        grad = torch.empty_like(out)
        nb = 1024
        part = torch.zeros(nb, dtype=torch.float64, device="cuda")
        _call("opy_mse3_f32", _hip.ptr(out), _hip.ptr(tgt), _hip.ptr(grad), N * H * W, self.n_out, wts[0], wts[1],
              wts[2], _hip.ptr(part), nb, _hip.stream())
        return part, grad

    def train_step(self, x, tgt, lr=1e-3, wts=(1.0, 1.0, 1.0)):
        out, acts = self.forward(x, keep=True)
        part, g = self.loss_grad(out, tgt, wts)
        self.backward(acts, g)
        self.adam(lr)
        return part


# -- encoding between images / solutions and network tensors --------------------
def encode_inputs(images, lams):
    """list of (H, W) host images (same shape) -> (N, H, W, 32) device tensor."""
    N = len(images)
    H, W = images[0].shape
    x = torch.zeros((N, H, W, CH), dtype=torch.float32, device="cuda")
    f = torch.as_tensor(np.stack(images).astype(np.float32)).cuda()
    x[..., 0] = f - 0.5
    x[..., 1] = torch.as_tensor(np.log(np.asarray(lams, dtype=np.float32)) / 3.0).cuda()[:, None, None]
    return x


def encode_targets(images, xs, ys):
    """Solutions (x (H*W,), y (H*W*2,)) -> (N, H, W, 32) target tensor (x - f, y_vertical, y_horizontal)."""
    N = len(images)
    H, W = images[0].shape
    t = np.zeros((N, H, W, 3), dtype=np.float32)
    for i in range(N):
        t[i, ..., 0] = xs[i].reshape(H, W) - images[i]
        t[i, ..., 1:3] = ys[i].reshape(H, W, 2)
    out = torch.zeros((N, H, W, CH), dtype=torch.float32, device="cuda")
    out[..., :3] = torch.as_tensor(t).cuda()
    return out


def _rows(t, off, stride, n, length):
    """n segments of `length` elements of the flat tensor t, `stride` apart from `off`, as an (n, length) view."""
    return t.as_strided((n, length), (stride, 1), storage_offset=t.storage_offset() + off)


class WarmStart(object):
    """Predicts packed (x0, y0) start vectors for a TVBatch of C=1 instances."""

    def __init__(self, net, tvw=1.0, chunk=16):
        self.net = net
        self.tvw = tvw
        self.chunk = chunk
        self._inp = {}  # input tensors by shape: channels 2.. stay zero, 0 and 1 are rewritten per chunk

    def predict_into(self, batch, lams, fused=True):
        """Writes the prediction into batch.x[0] / batch.y[0] (the start buffers).

        Instances of one shape go through the network in chunks.  A chunk
        whose instances sit at equal spacing in the packed buffers (consecutive
        instances of one shape) of an FP32 batch goes through
        ConvNet.forward_packed, whose first and last layers read the packed
        images and write the packed starts themselves (fused=False: through
        the encoded 32-channel tensors, as other chunks; the same values).
        Other chunks are read and written through one strided view per array,
        or instance by instance."""
        assert batch.C == 1
        f_all = batch.f
        groups = {}
        for b, shp in enumerate(batch.shapes):
            groups.setdefault(tuple(shp), []).append(b)
        for (H, W), idx in groups.items():
            HW = H * W
            for k in range(0, len(idx), self.chunk):
                sel = idx[k:k + self.chunk]
                n = len(sel)
                offs = [batch.offs[b] for b in sel]
                steps = {o1 - o0 for o0, o1 in zip(offs, offs[1:])}
                stride = steps.pop() if len(steps) == 1 else (HW if n == 1 else None)
                loglam = torch.as_tensor(np.log(np.asarray([lams[b] for b in sel], dtype=np.float64)) / 3.0,
                                         dtype=torch.float32).cuda()
                if fused and stride is not None and batch.dtype == torch.float32:
                    self.net.forward_packed(_hip.ptr(f_all[offs[0]:]), stride, loglam, n, H, W,
                                            _hip.ptr(batch.x[0][offs[0]:]), _hip.ptr(batch.y[0][2 * offs[0]:]))
                    self._project(batch, sel, offs, stride, HW)
                    continue
                inp = self._inp.get((n, H, W))
                if inp is None:
                    inp = self._inp[(n, H, W)] = torch.zeros((n, H, W, CH), dtype=torch.float32, device="cuda")
                inp[..., 1] = loglam[:, None, None]
                if stride is not None:
                    fv = _rows(f_all, offs[0], stride, n, HW)
                    inp[..., 0] = fv.view(n, H, W).float() - 0.5
                else:
                    for j, off in enumerate(offs):
                        inp[j, ..., 0] = f_all[off:off + HW].view(H, W).float() - 0.5
                out = self.net.forward(inp, bf16_acts=True)
                xv = (inp[..., 0] + 0.5 + out[..., 0]).reshape(n, HW).to(batch.dtype)
                yv = out[..., 1:3].reshape(n, 2 * HW).to(batch.dtype)
                if stride is not None:
                    _rows(batch.x[0], offs[0], stride, n, HW).copy_(xv)
                    _rows(batch.y[0], 2 * offs[0], 2 * stride, n, 2 * HW).copy_(yv)
                else:
                    for j, off in enumerate(offs):
                        batch.x[0][off:off + HW].copy_(xv[j])
                        batch.y[0][2 * off:2 * off + 2 * HW].copy_(yv[j])
                self._project(batch, sel, offs, stride, HW)

    def _project(self, batch, sel, offs, stride, HW):
        """Feasible dual start: every pixel's pair of y onto |y| <= tvw of its
        own instance (batch.tvw); the network predicts for self.tvw."""
        n = len(offs)
        radii = [float(batch.tvw[b]) for b in sel]
        if stride == HW and len(set(radii)) == 1:
            _hip.call("l1norms_proj", batch.dtype, _hip.ptr(batch.y[0][2 * offs[0]:2 * offs[0] + 2 * n * HW]),
                      n * HW, 1, 2, radii[0], 0, _hip.stream())
        else:
            for off, rad in zip(offs, radii):
                _hip.call("l1norms_proj", batch.dtype, _hip.ptr(batch.y[0][2 * off:2 * off + 2 * HW]), HW,
                          1, 2, rad, 0, _hip.stream())
