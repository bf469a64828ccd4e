"""The C-ABI library loads and exports every symbol include/*.h declares
(no GPU calls)."""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"\b(opy_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_symbols():
    names = declared_symbols()
    assert "opy_last_error" in names and "opy_grad_f32" in names
    assert len(names) > 50


def test_library_exports_all_declared_symbols():
    from opymize_amd import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.fail("libopyhip.so is not built (run python -m opymize_amd.csrc.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", _hip.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (opy_[a-z0-9_]+)", out.stdout))
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, "declared but not exported: %s" % missing
    lib = _hip.load()
    for n in declared_symbols():
        getattr(lib, n)  # dlsym only
    # every binding of the Python side is declared in the header
    undeclared = [n for n in _hip.exported_names() if n not in declared_symbols()]
    assert not undeclared, undeclared


def test_abi_version_and_error_string():
    from opymize_amd import _hip
    lib = _hip.load()
    assert lib.opy_abi_version() >= 1
    assert isinstance(lib.opy_last_error(), bytes)


def test_invalid_arguments_fail_without_gpu_work():
    """Argument validation happens on the host before any launch."""
    import ctypes
    from opymize_amd import _hip
    lib = _hip.load()
    g = _hip.make_grid((4, 4), 1, (1.0, 1.0))
    g.ndims = 9  # out of range
    rc = lib.opy_grad_f32(None, None, ctypes.byref(g), None, 0, 0, None)
    assert rc == 1 and b"ndims" in lib.opy_last_error()
    rc = lib.opy_l1norms_proj_f32(None, 10, 3, 3, 1.0, 1, None)
    assert rc == 1
    rc = lib.opy_quadepi_proj_f64(None, 10, 99, 1.0, None, float("inf"), None)
    assert rc == 1 and b"M=99" in lib.opy_last_error()


def test_tv_entry_points_refuse_c4_with_16_rows():
    """The fused TV entry points check their configuration before anything
    else: C=4 with 16 rows per thread is refused (no launch, no GPU)."""
    import ctypes
    from opymize_amd import _hip
    from opymize_amd.solvers.tv import TVCfg
    lib = _hip.load()
    c = TVCfg()
    c.C, c.tw, c.rpt, c.steps, c.h0, c.h1 = 4, 64, 16, 0, 1.0, 1.0
    for name in ("opy_tv_primal_f32", "opy_tv_primal_f64"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        rc = fn(None, None, None, None, None, None, None, None, None, ctypes.byref(c), 0, None)
        assert rc == 1 and b"register budget" in lib.opy_last_error()
    c.rpt = 12  # accepted by the check, then refused for the null pointers
    rc = lib.opy_tv_primal_f32(None, None, None, None, None, None, None, None, None, ctypes.byref(c), 0, None)
    assert rc == 1 and b"null pointer" in lib.opy_last_error()
