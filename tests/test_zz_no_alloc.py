"""The allocation-free residual entry point against the one with a work
buffer: opy_pdhg_residual2_noalloc_f32 must return the same two FP64 sums as
opy_pdhg_residual2_f32 (same kernels, same block count, so bitwise equal),
on two calls in a row, the second on a smaller input, with per-element and
device-held steps.  Both are called through the C-ABI."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bind(lib):
    fn = lib.opy_pdhg_residual2_noalloc_f32
    P, D, L = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64
    fn.argtypes = [P, P, P, D, P, D, D, P, L, P, P]
    fn.restype = ctypes.c_int
    return fn


def _case(n, seed, step_kind):
    rng = np.random.default_rng(seed)
    t = {k: torch.as_tensor(rng.standard_normal(n).astype(np.float32), device="cuda")
         for k in ("z", "zn", "g", "gn")}
    if step_kind == "array":
        t["a"] = torch.as_tensor((0.5 + rng.random(n)).astype(np.float32), device="cuda")
    else:
        t["a"] = torch.tensor([0.37], dtype=torch.float64, device="cuda")
    return t


@pytest.mark.parametrize("step_kind", ["array", "device"])
def test_noalloc_residual_matches_work_buffer_version(gpu, step_kind):
    from opymize_amd import _hip
    lib = _hip.load()
    noalloc = _bind(lib)
    # first call: more elements than 1024 blocks of 256 threads cover (grid-stride path);
    # second: a smaller input, so stale partials of the first call would show
    for n, seed in ((1024 * 256 * 3 + 17, 1), (5000, 2)):
        print("no-alloc residual: n=%d seed=%d step=%s" % (n, seed, step_kind), flush=True)
        t = _case(n, seed, step_kind)
        st = _hip.make_step(t["a"], torch.float32)
        ref = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
        out = torch.full((2,), -2.0, dtype=torch.float64, device="cuda")
        work = _hip.Workspace.get()
        s = _hip.stream()
        args = (_hip.ptr(t["z"]), _hip.ptr(t["zn"]), 0.75, _hip.ptr(t["g"]), 0.5, 1.5, _hip.ptr(t["gn"]), n)
        _hip.call("pdhg_residual2", torch.float32, ctypes.byref(st), *args, _hip.ptr(work), _hip.ptr(ref), s)
        rc = noalloc(ctypes.byref(st), *args, _hip.ptr(out), s)
        assert rc == 0, lib.opy_last_error()
        torch.cuda.synchronize()
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert np.array_equal(r, o), (r, o)
        # and against a float64 restatement of the two sums (pdhg.py:74-78)
        h = {k: v.cpu().numpy().astype(np.float64) for k, v in t.items()}
        a = h["a"] if step_kind == "array" else float(h["a"][0])
        d = (h["z"] - h["zn"]) / a - 0.75 * h["g"]
        want = np.array([((d + 0.5 * h["gn"]) ** 2).sum(), ((d + 1.5 * h["gn"]) ** 2).sum()])
        np.testing.assert_allclose(o, want, rtol=1e-5)


def test_noalloc_rejects_negative_size(gpu):
    from opymize_amd import _hip
    lib = _hip.load()
    noalloc = _bind(lib)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    rc = noalloc(None, None, None, 1.0, None, 0.0, 0.0, None, -5, _hip.ptr(out), _hip.stream())
    assert rc == 1 and b"negative" in lib.opy_last_error()
