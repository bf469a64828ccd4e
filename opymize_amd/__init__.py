"""MI355X-native implementation of Opymize's PDHG hot path.

Core abstractions with the reference's API (reference: opymize/__init__.py):
``Variable`` packs named sub-arrays into one flat C-order vector,
``Operator``/``LinOp`` are (matrix-free) maps evaluated by HIP kernels, and
``Functional`` exposes ``(value, infeasibility)`` evaluation and ``prox``.

Device arrays are torch ROCm tensors.  Host (numpy) arrays are accepted for
convenience: they are staged to the GPU, processed by the same kernels and
copied back, so every evaluation runs on the GPU.  Without a GPU the calls
fail loudly.
"""
import numpy as np

from opymize_amd import _hip


def _is_dev(x):
    return _hip.is_device_tensor(x)


class Variable(object):
    """Named blocks of one flat vector (reference: opymize/__init__.py:6-46)."""

    def __init__(self, *args):
        if len(args) == 1 and isinstance(args[0], (int, np.integer)):
            args = [(int(args[0]),)]
        if not isinstance(args[0][0], str):
            args = [("v%d" % i, a) for i, a in enumerate(args)]
        self._descr = {}
        self._args = args
        self.size = 0
        for name, dim in args:
            if isinstance(dim, (int, np.integer)):
                dim = (int(dim),)
            dim = tuple(int(d) for d in dim)
            n = int(np.prod(dim))
            self._descr[name] = {"name": name, "offset": self.size, "shape": dim, "size": n}
            self.size += n

    def vars(self, data=None, named=False):
        if data is None:
            if named:
                return self._descr
            return [self._descr[a[0]] for a in self._args]
        result = {} if named else [None] * len(self._args)
        for i, a in enumerate(self._args):
            d = self._descr[a[0]]
            start, end = d["offset"], d["offset"] + d["size"]
            result[d["name"] if named else i] = data[start:end].reshape(d["shape"])
        return result

    def __getitem__(self, idx):
        if isinstance(idx, str):
            return self._descr[idx]
        return self._descr[self._args[idx][0]]

    def new(self, dtype=np.float64):
        """Zero host vector (reference semantics)."""
        return np.zeros((self.size,), order="C", dtype=dtype)

    def new_device(self, dtype=np.float64):
        """Zero device vector (torch ROCm tensor)."""
        import torch
        _hip.require_device()
        return torch.zeros((self.size,), dtype=_hip.torch_dtype(dtype), device="cuda")


def _stage(x, y, add):
    """Upload host operands; returns (xd, yd, writeback_target)."""
    import torch
    _hip.require_device()
    xa = np.asarray(x)
    dt = torch.float32 if xa.dtype == np.float32 else torch.float64
    xd = _hip.to_device(xa.ravel(), dt)
    if y is None:
        return xd, None, x
    ya = np.asarray(y)
    yd = _hip.to_device(ya.ravel(), dt) if add else torch.empty(ya.size, dtype=dt, device="cuda")
    return xd, yd, y


def _writeback(dst, src_dev):
    out = src_dev.cpu().numpy()
    dst_flat = dst.reshape(-1) if isinstance(dst, np.ndarray) else dst
    dst_flat[...] = out.astype(dst_flat.dtype, copy=False)


class Operator(object):
    """A (nonlinear) map T evaluated on the GPU (reference: opymize/__init__.py:48-84)."""

    def __init__(self):
        self.x = None  # Variable, domain
        self.y = None  # Variable, range

    def prepare_gpu(self, type_t="double"):
        """Kernels are compiled ahead of time; this only validates the precision."""
        _hip.torch_dtype(type_t)

    def _call_gpu(self, x, y=None, add=False, jacobian=False):
        raise Exception("%s has no device implementation" % type(self).__name__)

    def __call__(self, x, y=None, add=False, jacobian=False):
        """Evaluate T(x); written to x unless y is given; add accumulates."""
        if _is_dev(x):
            return self._call_gpu(x, y=y, add=add, jacobian=jacobian)
        xd, yd, dst = _stage(x, y, add)
        res = self._call_gpu(xd, y=yd, add=add, jacobian=jacobian)
        _writeback(dst, xd if yd is None else yd)
        return res


class LinOp(Operator):
    """Matrix-free linear operator A (reference: opymize/__init__.py:86-142)."""

    def __init__(self):
        Operator.__init__(self)
        self.adjoint = None

    def _call_gpu(self, x, y=None, add=False):
        raise Exception("%s has no device implementation" % type(self).__name__)

    def __call__(self, x, y=None, add=False, jacobian=False):
        if _is_dev(x):
            self._call_gpu(x, y=y, add=add)
            return self if jacobian else None
        xd, yd, dst = _stage(x, y, add)
        self._call_gpu(xd, y=yd, add=add)
        _writeback(dst, xd if yd is None else yd)
        return self if jacobian else None

    def _rowwise_gpu(self, y, p=1, add=False):
        raise Exception("%s has no row-wise l^p norms" % type(self).__name__)

    def rowwise_lp(self, y, p=1, add=False):
        """y_i (+)= sum_j |A_ij|^p."""
        if _is_dev(y):
            return self._rowwise_gpu(y, p=p, add=add)
        import torch
        _hip.require_device()
        ya = np.asarray(y)
        dt = torch.float32 if ya.dtype == np.float32 else torch.float64
        yd = _hip.to_device(ya.ravel(), dt)
        self._rowwise_gpu(yd, p=p, add=add)
        _writeback(y, yd)


class Functional(object):
    """F with (value, infeasibility) evaluation and prox
    (reference: opymize/__init__.py:144-179).

    Evaluation runs as FP64 reductions on the GPU: ``_gpu_eval(x, out3)``
    enqueues reductions into a device double[3] (sum, max, max) and
    ``_gpu_finish(v)`` turns those numbers into ``(value, infeas)``.
    """

    def __init__(self):
        self.x = None
        self.conj = None

    def nslots(self):
        """Number of (sum, max, max) FP64 result triples _gpu_eval writes."""
        return 1

    # Width of the blocks of entries that the prox couples (a projection onto a
    # non-separable set: an epigraph, a simplex), or None for a prox that acts
    # on each entry alone or has a step-independent exact form per entry.
    precond_block = None

    def precond_blocks(self):
        """(offset, nblocks, width) ranges of x on which preconditioned steps
        must be constant for prox(tau) to be the exact prox."""
        w = self.precond_block
        if not w or self.x is None:
            return []
        return [(0, self.x.size // w, w)]

    def _gpu_eval(self, x, out3):
        raise Exception("%s cannot be evaluated" % type(self).__name__)

    def _gpu_finish(self, v):
        raise Exception("%s cannot be evaluated" % type(self).__name__)

    def __call__(self, x, grad=False):
        if grad:
            raise NotImplementedError("gradients of functionals belong to the Newton solvers, "
                                      "which are outside the PDHG hot path")
        import torch
        if not _is_dev(x):
            xa = np.asarray(x)
            dt = torch.float32 if xa.dtype == np.float32 else torch.float64
            x = _hip.to_device(xa.ravel(), dt)
        out = torch.zeros(3 * self.nslots(), dtype=torch.float64, device=x.device)
        self._gpu_eval(x.reshape(-1), out)
        return self._gpu_finish(out.cpu().numpy())

    def prox(self, tau):
        raise Exception("%s has no prox operator" % type(self).__name__)
