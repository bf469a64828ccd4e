"""Ahead-of-time build of libopyhip.so for MI355X (gfx950).

Every translation unit under this directory is compiled by hipcc into an
object file under ``_build/`` and linked into ``opymize_amd/libopyhip.so``.
No JIT: the shared library travels with the source tree.

Usage: ``python -m opymize_amd.csrc.build [-j N] [--force]``
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(PKG, "libopyhip.so")
ARCH = os.environ.get("OPY_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17",
    "--offload-arch=" + ARCH,
    # code-object v5 loads on both the system ROCm runtime and torch's bundled one
    "-mcode-object-version=5",
    # never emit packed FP32 arithmetic (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def sources():
    out = []
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(HERE, f))
    return out


def headers():
    hs = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    hs += [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".h")]
    return hs


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def _stale(src, obj, hdr_mtime):
    if not os.path.exists(obj):
        return True
    m = os.path.getmtime(obj)
    return os.path.getmtime(src) > m or hdr_mtime > m


def compile_one(src):
    obj = _obj(src)
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(jobs=None, force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    hdr_mtime = max(os.path.getmtime(h) for h in headers())
    todo = [s for s in srcs if force or _stale(s, _obj(s), hdr_mtime)]
    jobs = jobs or min(8, os.cpu_count() or 1)
    if todo:
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            for obj in ex.map(compile_one, todo):
                if verbose:
                    print("compiled", os.path.basename(obj), flush=True)
    objs = [_obj(s) for s in srcs]
    lib_stale = force or bool(todo) or not os.path.exists(LIB) or \
        max(os.path.getmtime(o) for o in objs) > os.path.getmtime(LIB)
    if lib_stale:
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
        os.replace(tmp, LIB)
        if verbose:
            print("linked", LIB, flush=True)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args(argv)
    build(jobs=a.j, force=a.force, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
