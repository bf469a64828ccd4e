"""ctypes boundary to libopyhip.so (the HIP kernels of this package).

Torch is imported first so that the process has exactly one HIP runtime:
torch's bundled ``libamdhip64.so.7`` satisfies the library's dependency by
soname.  Torch tensors serve only as device buffers; every computation on them
in this package is a kernel of ``libopyhip.so``.

If the library is missing the import of any operator fails loudly: there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopyhip.so")

OPY_MAX_DIMS = 6
OPY_RED_BLOCKS = 1024
OPY_QUADEPI_MAXM = 16
SCHEME_FORWARD = 0
SCHEME_CENTERED = 1

FN_SSD = 1
FN_L1DATA = 2
FN_L1DATA_CONJ = 3
FN_NORMS = 4
FN_DOT = 5
FN_INDICATOR = 6
FN_MASKED_AFFINE = 7
FN_POS = 8
FN_NEG = 9
FN_QUADEPI_IND = 10
FN_QUADEPI_SUPP = 11
FN_HUBER = 12
FN_TRUNCQUADEPI = 13
FN_EPIGRAPH_IND = 14
FN_EPIGRAPH_SUPP = 15
FN_BOX_IND = 16
FN_BOX_SUPP = 17
FN_SIMPLEX_IND = 18
FN_SIMPLEX_SUPP = 19


class Grid(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int32), ("nchannels", ctypes.c_int32),
                ("dims", ctypes.c_int64 * OPY_MAX_DIMS), ("h", ctypes.c_double * OPY_MAX_DIMS)]


class Step(ctypes.Structure):
    _fields_ = [("arr", ctypes.c_void_p), ("dev", ctypes.c_void_p), ("host", ctypes.c_double)]


class AdaptState(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double), ("sigma", ctypes.c_double), ("alpha", ctypes.c_double),
                ("resp", ctypes.c_double), ("resd", ctypes.c_double),
                ("epsp", ctypes.c_double), ("epsd", ctypes.c_double),
                ("sums", ctypes.c_double * 4), ("nchanges", ctypes.c_int32), ("pad", ctypes.c_int32)]


ADAPT_STATE_DOUBLES = ctypes.sizeof(AdaptState) // 8
ADAPT_FIELD = {name: getattr(AdaptState, name).offset // 8 for name, _ in AdaptState._fields_}


class AdaptParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("eta", "Delta", "sqrt_n", "sqrt_m", "eps_absp", "eps_relp", "eps_absd", "eps_reld")]


class FnParams(ctypes.Structure):
    _fields_ = [("f", ctypes.c_void_p), ("vol", ctypes.c_void_p), ("mask", ctypes.c_void_p),
                ("c", ctypes.c_void_p), ("cs", ctypes.c_double), ("lbd", ctypes.c_double),
                ("m1", ctypes.c_int32), ("m2", ctypes.c_int32), ("matrixnorm", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("qa", ctypes.c_double), ("qb", ctypes.c_void_p),
                ("qc", ctypes.c_void_p), ("ei", ctypes.c_void_p), ("ej", ctypes.c_void_p),
                ("ev", ctypes.c_void_p), ("eb", ctypes.c_void_p), ("eq", ctypes.c_void_p),
                ("eq_off", ctypes.c_void_p), ("checks", ctypes.c_void_p), ("nfuns", ctypes.c_int32),
                ("npoints", ctypes.c_int32), ("nsub", ctypes.c_int32), ("epad", ctypes.c_int32)]


_P, _I, _L, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double

# name -> argtypes (without the typed suffix); every function returns int
_SIGS = {
    "grad": [_P, _P, _P, _P, _I, _I, _P],
    "div": [_P, _P, _P, _P, _I, _I, _P],
    "diff_rowwise": [_P, _P, _P, _I, _I, _D, _I, _P],
    "laplacian": [_P, _P, _P, _I, _I, _I, _P],
    "laplacian_rowwise": [_P, _P, _I, _I, _D, _I, _P],
    "matmul": [_P, _P, _P, _L, _I, _I, _I, _I, _P],
    "matmul_r": [_P, _P, _P, _L, _I, _I, _I, _I, _P],
    "diag_matmul_r": [_P, _P, _L, _I, _P, _I, _P],
    "indexedmult": [_P, _P, _P, _P, _L, _I, _I, _I, _I, _I, _I, _P],
    "tangled_matmul_r": [_P, _P, _P, _L, _I, _I, _I, _I, _I, _I, _P],
    "matmul_r_batched": [_P, _P, _P, _L, _I, _I, _I, _I, _I, _P],
    "csr_spmv": [_L, _P, _P, _P, _P, _P, _I, _P],
    "csr_rowwise": [_L, _P, _P, _D, _P, _I, _P],
    "fill": [_P, _L, _D, _P],
    "copy": [_P, _P, _L, _P],
    "scale": [_P, _P, _L, _P, _D, _I, _P],
    "tile": [_P, _L, _P, _L, _L, _I, _P],
    "abs": [_P, _P, _L, _P],
    "recip_gt": [_P, _L, _D, _P],
    "nihil": [_P, _P, _L, _P, _I, _P],
    "shiftscale": [_P, _P, _L, _P, _D, _P, _D, _P, _D, _I, _P],
    "ssd_prox": [_P, _P, _L, _I, _P, _P, _P, _P, _I, _P],
    "l1data_prox": [_P, _P, _L, _I, _P, _P, _I, _P, _I, _P],
    "signproj": [_P, _P, _L, _I, _I, _P],
    "const": [_P, _L, _P, _I, _P],
    "constrain": [_P, _P, _L, _L, _P, _P, _I, _P],
    "l1norms_proj": [_P, _L, _I, _I, _D, _I, _P],
    "quadepi_proj": [_P, _L, _I, _D, _P, _D, _P],
    "epigraph_proj": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _P],
    "pdhg_take_step": [_P, _P, _L, _I, _P, _P, _P],
    "pdhg_dual_step": [_P, _P, _L, _D, _P, _P, _P, _P],
    "pdhg_residual2": [_P, _P, _P, _D, _P, _D, _D, _P, _L, _P, _P, _P],
    "fn_eval": [_I, _P, _L, _I, _P, _P, _P, _I, _P],
    "clip": [_P, _P, _L, _D, _D, _P],
    "simplex_proj": [_P, _L, _I, _P],
    "block_min": [_P, _L, _I, _P],
    "dot": [_P, _P, _L, _P, _P, _P],
    "axpby": [_P, _P, _P, _L, _D, _D, _P],
    "signproj_jac": [_P, _P, _P, _L, _I, _I, _P],
    "l12proj_jac": [_P, _P, _P, _L, _I, _D, _I, _P],
}
# untyped entry points (no _f32/_f64 suffix)
_SIGS_PLAIN = {
    "opy_pdhg_adapt": [_P, _P, _P],
    "opy_abi_version": [],
    "opy_device_sync": [],
}

_lib = None


def exported_names():
    """Every symbol the Python side binds (used by the ABI export test)."""
    names = []
    for k in _SIGS:
        names += ["opy_%s_f32" % k, "opy_%s_f64" % k]
    names += list(_SIGS_PLAIN)
    names.append("opy_last_error")
    return names


def load(path=None):
    """Load libopyhip.so and bind prototypes (no GPU is touched)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError("opymize_amd: %s is not built; run `python -m opymize_amd.csrc.build`" % p)
    lib = ctypes.CDLL(p)
    lib.opy_last_error.restype = ctypes.c_char_p
    lib.opy_last_error.argtypes = []
    for k, args in _SIGS.items():
        for sfx in ("_f32", "_f64"):
            fn = getattr(lib, "opy_%s%s" % (k, sfx))
            fn.restype = ctypes.c_int
            fn.argtypes = args
    for k, args in _SIGS_PLAIN.items():
        fn = getattr(lib, k)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class HipError(RuntimeError):
    pass


def _check(rc, name):
    if rc != 0:
        msg = _lib.opy_last_error().decode(errors="replace")
        raise HipError("%s failed (status %d): %s" % (name, rc, msg))


def sfx(dtype):
    if dtype == torch.float32:
        return "_f32"
    if dtype == torch.float64:
        return "_f64"
    raise TypeError("opymize_amd kernels support float32 and float64, got %s" % dtype)


def call(name, dtype, *args):
    """Call opy_<name>_<f32|f64>(*args) and raise HipError on failure."""
    lib = load()
    fname = "opy_%s%s" % (name, sfx(dtype))
    _check(getattr(lib, fname)(*args), fname)


def call_plain(name, *args):
    lib = load()
    _check(getattr(lib, name)(*args), name)


def stream():
    """Handle of torch's current stream on the current device."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_device():
    if not torch.cuda.is_available():
        raise RuntimeError("opymize_amd: no ROCm GPU is visible; the PDHG hot path runs only on the GPU")


def is_device_tensor(x):
    return isinstance(x, torch.Tensor) and x.is_cuda


def ptr(t, dtype=None, numel=None):
    """Device pointer of a contiguous CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a ROCm device tensor, got %r" % type(t))
    if not t.is_contiguous():
        raise ValueError("device tensor must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError("expected dtype %s, got %s" % (dtype, t.dtype))
    if numel is not None and t.numel() < numel:
        raise ValueError("device tensor has %d elements, need %d" % (t.numel(), numel))
    return ctypes.c_void_p(t.data_ptr())


def addr(t, offset_elems=0):
    """Raw device address (int) of t[offset_elems] as c_void_p."""
    return ctypes.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def to_device(a, dtype):
    """Upload a host array / scalar to a new contiguous device tensor."""
    require_device()
    t = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype)
    return t.to(device="cuda", non_blocking=False).contiguous()


def make_grid(imagedims, nchannels, imageh):
    dims = [int(d) for d in imagedims]
    if not 1 <= len(dims) <= OPY_MAX_DIMS:
        raise ValueError("1 to %d image dimensions are supported, got %d" % (OPY_MAX_DIMS, len(dims)))
    g = Grid()
    g.ndims = len(dims)
    g.nchannels = int(nchannels)
    for t in range(OPY_MAX_DIMS):
        g.dims[t] = dims[t] if t < len(dims) else 1
        g.h[t] = float(imageh[t]) if t < len(dims) else 1.0
    return g


def make_step(step, dtype):
    """opy_step_t from a float, a device scalar (float64 tensor of one element
    or (tensor, index)), or a per-element device array."""
    s = Step()
    s.arr = None
    s.dev = None
    s.host = 0.0
    if isinstance(step, tuple):  # (float64 device tensor, element index)
        t, idx = step
        s.dev = t.data_ptr() + 8 * int(idx)
    elif isinstance(step, torch.Tensor):
        if step.dtype == torch.float64 and step.numel() == 1 and dtype != torch.float64:
            s.dev = step.data_ptr()
        elif step.numel() == 1 and step.dtype == torch.float64:
            s.dev = step.data_ptr()
        else:
            if step.dtype != dtype:
                raise TypeError("step array dtype %s != %s" % (step.dtype, dtype))
            s.arr = step.data_ptr()
    else:
        s.host = float(step)
    return s


def torch_dtype(type_t):
    if type_t in ("double", np.float64, torch.float64):
        return torch.float64
    if type_t in ("float", np.float32, torch.float32):
        return torch.float32
    raise ValueError("unknown precision %r (use 'double' or 'float')" % (type_t,))


def np_dtype(dtype):
    return np.float64 if dtype == torch.float64 else np.float32


class Workspace:
    """Per-device scratch for two-stage FP64 reductions (allocated once)."""
    _bufs = {}

    @classmethod
    def get(cls):
        dev = torch.cuda.current_device()
        b = cls._bufs.get(dev)
        if b is None:
            b = torch.zeros(3 * OPY_RED_BLOCKS, dtype=torch.float64, device="cuda")
            cls._bufs[dev] = b
        return b
