"""Timing of the FP32 per-point matrix products (MatrixMult, MatrixMultR) on
the matrix cores, for profiles: each shape and size runs a few warm-up
applies and then `reps` timed ones (HIP events on the launch stream), and
prints the time and the algorithmic bandwidth, N * (K + J) * 4 bytes per
apply; after each shape, the library's copy kernel moving the same bytes.

Usage: python -m opymize_amd.tools.einsum_bench [--sizes N ...] [--reps R]
"""
import argparse

import numpy as np

SHAPES = ((16, 16), (15, 30), (30, 15))
HBM_PEAK_GBS = 8000.0


def main(argv=None):
    import torch
    from opymize_amd import _hip
    from opymize_amd.linear import MatrixMult, MatrixMultR
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[512 * 512, 4096 * 1024])
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    rng = np.random.default_rng(0)
    for N in a.sizes:
        for (K, J) in SHAPES:
            for cls in (MatrixMult, MatrixMultR):
                # MatrixMult: A is J x K (y = A x, label-major); MatrixMultR: A is K x J (y = x A, point-major)
                A = rng.standard_normal((J, K) if cls is MatrixMult else (K, J))
                op = cls(N, A)
                x = torch.randn(op.x.size, device="cuda", dtype=torch.float32)
                y = torch.zeros(op.y.size, device="cuda", dtype=torch.float32)
                for _ in range(3):
                    op(x, y)
                torch.cuda.synchronize()
                st = torch.cuda.current_stream()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                for _ in range(a.reps):
                    op(x, y)
                e1.record(st)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / a.reps
                gbs = N * (K + J) * 4 / (ms * 1e-3) / 1e9
                print("%-11s N=%9d K=%2d J=%2d: %8.1f us  %6.0f GB/s  %.2f of peak"
                      % (cls.__name__, N, K, J, 1e3 * ms, gbs, gbs / HBM_PEAK_GBS), flush=True)
                del x, y, op
            # the library's copy kernel moving the same bytes (N (K + J) / 2 floats read and
            # written): what a launch of this size reaches without arithmetic
            n = N * (K + J) // 2
            a_, b_ = torch.ones(n, device="cuda"), torch.empty(n, device="cuda")
            cp = lambda: _hip.call("copy", a_.dtype, _hip.ptr(a_), _hip.ptr(b_), n, _hip.stream())
            for _ in range(3):
                cp()
            torch.cuda.synchronize()
            e0.record(st)
            for _ in range(a.reps):
                cp()
            e1.record(st)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / a.reps
            gbs = 2 * n * 4 / (ms * 1e-3) / 1e9
            print("%-11s N=%9d K=%2d J=%2d: %8.1f us  %6.0f GB/s  %.2f of peak"
                  % ("copy", N, K, J, 1e3 * ms, gbs, gbs / HBM_PEAK_GBS), flush=True)
            del a_, b_


if __name__ == "__main__":
    main()
