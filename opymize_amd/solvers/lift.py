"""Batched fused PDHG for the multi-label lifting model.

Many instances of opymize_amd.lifting.sublabel_tv -- images of uneven size
and uneven label count -- iterate together: two kernel launches per
iteration over the tiles of the still-running instances (csrc/lift.hip),
blocks of iterations captured in HIP graphs, an FP64 gap per instance on the
device every check, and a device-side list of the running instances.  The
iterates equal those of PDHG(G, F, A).solve(steps="precond") of each
instance's model (reference pdhg.py:229-315 with its steps "precond",
pdhg.py:189-200, made constant on the coupled blocks), up to rounding.
"""
import ctypes

import numpy as np
import torch

from opymize_amd import _hip
from opymize_amd.lifting import labels

TILE_W, TILE_H = 64, 4
PARTIALS_PER_TILE = 8
CHECK_DOUBLES = 9  # opy_tv_check_t: 8 doubles + 2 int32
ALIGN = 64


class LiftInst(ctypes.Structure):
    _fields_ = [("off", ctypes.c_int64), ("offk", ctypes.c_int64), ("H", ctypes.c_int32), ("W", ctypes.c_int32),
                ("K", ctypes.c_int32), ("tile0", ctypes.c_int32), ("ntiles", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("lam", ctypes.c_double), ("g0", ctypes.c_double),
                ("h", ctypes.c_double), ("tau_r", ctypes.c_double), ("omega", ctypes.c_double)]


P, I, L, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
_SIGS = {
    "opy_lift_primal": [P, P, P, P, P, P, P, P, P, P, I, I, P],
    "opy_lift_dual": [P, P, P, P, P, P, P, P, P, I, I, P],
    "opy_lift_gap": [P, P, P, P, P, P, P, P, P, I, P, I, P],
}
_PLAIN = {
    "opy_lift_check": [P, P, P, I, D, D, P, I, P],
    "opy_lift_compact": [P, P, I, P, P, P, P, I, P],
}
_bound = False


def _bind():
    global _bound
    if _bound:
        return
    lib = _hip.load()
    for k, a in _SIGS.items():
        for s in ("_f32", "_f64"):
            fn = getattr(lib, k + s)
            fn.restype, fn.argtypes = ctypes.c_int, a
    for k, a in _PLAIN.items():
        fn = getattr(lib, k)
        fn.restype, fn.argtypes = ctypes.c_int, a
    _bound = True


def exported_names():
    return [k + s for k in _SIGS for s in ("_f32", "_f64")] + list(_PLAIN)


class LiftBatch(object):
    """images: list of 2-D arrays (H_b, W_b); L: labels per instance (scalar
    or list, 2..33); lam: TV weight per instance; lo, hi: label range per
    instance (default each image's min and max); omega: primal weight per
    instance, the diagonal steps become omega*T and Sigma/omega (the
    convergence condition |Sigma^1/2 K T^1/2| <= 1 does not change; 1 gives
    the reference's preconditioner)."""

    def __init__(self, images, L, lam, lo=None, hi=None, precision="float", omega=1.0):
        imgs = [np.asarray(im, dtype=np.float64) for im in images]
        if not imgs or any(im.ndim != 2 for im in imgs):
            raise ValueError("LiftBatch needs a non-empty list of 2-D images")
        if not all(np.isfinite(im).all() for im in imgs) or any(im.size == 0 for im in imgs):
            raise ValueError("images must be non-empty and finite")
        self.B = len(imgs)
        Ls = np.broadcast_to(np.asarray(L, dtype=np.int64), (self.B,)).copy()
        if np.any(Ls < 2) or np.any(Ls > 33):
            raise ValueError("labels per instance must lie in [2, 33] (at most 32 pieces)")
        lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (self.B,)).copy()
        if np.any(lam <= 0):
            raise ValueError("lam must be positive")
        omega = np.broadcast_to(np.asarray(omega, dtype=np.float64), (self.B,)).copy()
        if np.any(~(omega > 0)):
            raise ValueError("omega must be positive")
        self.omega = omega
        los = [im.min() if lo is None else np.broadcast_to(lo, (self.B,))[b] for b, im in enumerate(imgs)]
        his = [im.max() if hi is None else np.broadcast_to(hi, (self.B,))[b] for b, im in enumerate(imgs)]
        self.gam = [labels(Ls[b], los[b], his[b]) for b in range(self.B)]
        self.shapes = [im.shape for im in imgs]
        self.K = [int(l) - 1 for l in Ls]
        self.kmax = max(self.K)
        self.lam = lam
        self.dtype = _hip.torch_dtype(precision)
        _hip.require_device()
        _bind()
        insts = (LiftInst * self.B)()
        off = offk = t0 = 0
        self.offs, self.offks, self.ntiles = [], [], []
        for b, (H, W) in enumerate(self.shapes):
            N, k = H * W, self.K[b]
            g = self.gam[b]
            h = float(g[1] - g[0])
            nt = ((H + TILE_H - 1) // TILE_H) * ((W + TILE_W - 1) // TILE_W)
            rec = insts[b]
            rec.off, rec.offk, rec.H, rec.W, rec.K = off, offk, H, W, k
            rec.tile0, rec.ntiles = t0, nt
            rec.lam, rec.g0, rec.h = float(lam[b]), float(g[0]), h
            # column sums of |K| for r_l: |g_l| + h*l (a_i, i < l) + 1 (the pair's -r_l); the smallest step
            rec.tau_r = float(np.min(1.0 / (np.abs(g[:k]) + h * np.arange(k) + 1.0)))
            rec.omega = float(omega[b])
            self.offs.append(off)
            self.offks.append(offk)
            self.ntiles.append(nt)
            off += -(-N // ALIGN) * ALIGN
            offk += -(-N * k // ALIGN) * ALIGN
            t0 += nt
        self.total_tiles = t0
        self.n_pix, self.n_planes = off, offk
        self.insts = torch.as_tensor(np.frombuffer(bytes(insts), dtype=np.uint8).copy()).cuda()
        fpack = np.zeros(off)
        for im, o in zip(imgs, self.offs):
            fpack[o:o + im.size] = im.ravel()
        dt = self.dtype
        self.f = _hip.to_device(fpack, dt)
        self.U = torch.zeros(offk, dtype=dt, device="cuda")
        self.r = torch.zeros(offk, dtype=dt, device="cuda")
        self.Ub = torch.zeros(offk, dtype=dt, device="cuda")
        self.rb = torch.zeros(offk, dtype=dt, device="cuda")
        self.p = torch.zeros(2 * offk, dtype=dt, device="cuda")
        self.z = torch.zeros(2 * offk, dtype=dt, device="cuda")
        self.partials = torch.zeros(PARTIALS_PER_TILE * max(t0, 1), dtype=torch.float64, device="cuda")
        self.checks = torch.zeros(self.B * CHECK_DOUBLES, dtype=torch.float64, device="cuda")
        self.active_inst = torch.zeros(self.B, dtype=torch.int32, device="cuda")
        self.act_off = torch.zeros(self.B + 1, dtype=torch.int32, device="cuda")
        self.nact = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.iter_counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.grid = max(1, t0)

    # -- kernels ---------------------------------------------------------------
    def _sfx(self):
        return "_f32" if self.dtype == torch.float32 else "_f64"

    def _call(self, name, *args):
        lib = _hip.load()
        rc = getattr(lib, name)(*args)
        if rc != 0:
            raise _hip.HipError("%s failed: %s" % (name, lib.opy_last_error().decode()))

    def _iteration(self):
        Pt, s = _hip.ptr, _hip.stream()
        ctl = (Pt(self.insts), Pt(self.active_inst), Pt(self.act_off), Pt(self.nact), self.kmax, self.grid, s)
        self._call("opy_lift_primal" + self._sfx(), Pt(self.U), Pt(self.r), Pt(self.Ub), Pt(self.rb), Pt(self.p),
                   Pt(self.z), *ctl)
        self._call("opy_lift_dual" + self._sfx(), Pt(self.Ub), Pt(self.rb), Pt(self.p), Pt(self.z), Pt(self.f),
                   *ctl)

    def _check(self, tol, tol_infeas, iters):
        Pt, s = _hip.ptr, _hip.stream()
        self._call("opy_lift_gap" + self._sfx(), Pt(self.U), Pt(self.r), Pt(self.p), Pt(self.z), Pt(self.f),
                   Pt(self.insts), Pt(self.active_inst), Pt(self.act_off), Pt(self.nact), self.kmax,
                   Pt(self.partials), self.grid, s)
        self._call("opy_lift_check", Pt(self.insts), Pt(self.checks), Pt(self.partials), self.B, float(tol),
                   float(tol_infeas), Pt(self.iter_counter), int(iters), s)
        self._compact(iters)

    def _compact(self, iters):
        Pt = _hip.ptr
        self._call("opy_lift_compact", Pt(self.insts), Pt(self.checks), self.B, Pt(self.active_inst),
                   Pt(self.act_off), Pt(self.nact), Pt(self.iter_counter), int(iters), _hip.stream())

    def refresh_grid(self):
        """One block per running tile (one device read); returns the running count."""
        nrun = int(self.nact.item())
        tiles = int(self.act_off[nrun].item()) if nrun > 0 else 0
        self.grid = max(1, tiles)
        return nrun

    @staticmethod
    def _capture(fn):
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                fn()
        torch.cuda.current_stream().wait_stream(s)
        return g

    # -- public ----------------------------------------------------------------
    def reset(self):
        """x = y = 0 for every instance, all running."""
        for t in (self.U, self.r, self.Ub, self.rb, self.p, self.z):
            t.zero_()
        self.checks.zero_()
        self.iter_counter.zero_()
        self._compact(0)

    def set_start(self, xs=None, ys=None):
        """Start values per instance in the layout of solution() (x = (U, r),
        y = (p (N*k, 3), z (N*k, 2)); None keeps zeros); all instances running."""
        U = np.zeros(self.n_planes)
        r = np.zeros(self.n_planes)
        p = np.zeros(2 * self.n_planes)
        z = np.zeros(2 * self.n_planes)
        for b, (H, W) in enumerate(self.shapes):
            N, k, ok = H * W, self.K[b], self.offks[b]
            if xs is not None and xs[b] is not None:
                x = np.asarray(xs[b], dtype=np.float64).ravel()
                U[ok:ok + N * k] = x[:N * k].reshape(N, k).T.ravel()
                r[ok:ok + N * k] = x[N * k:2 * N * k].reshape(N, k).T.ravel()
            if ys is not None and ys[b] is not None:
                y = np.asarray(ys[b], dtype=np.float64).ravel()
                p3 = y[:3 * N * k].reshape(N, k, 3)
                zz = y[3 * N * k:5 * N * k].reshape(N, k, 2)
                p[2 * ok:2 * ok + 2 * N * k] = p3[:, :, :2].transpose(1, 2, 0).ravel()
                z[2 * ok:2 * ok + 2 * N * k] = zz.transpose(1, 2, 0).ravel()
        for t, v in ((self.U, U), (self.r, r), (self.p, p), (self.z, z)):
            t.copy_(torch.as_tensor(v, dtype=self.dtype))
        self.checks.zero_()
        self.iter_counter.zero_()
        self._compact(0)

    def solve(self, max_iters=100000, block_iters=100, check_every=10, tol=1e-5, tol_infeas=None, use_graph=True,
              callback=None):
        """Iterate from the current state (zeros after construction or reset())
        until every instance has |relgap| < tol and both infeasibilities
        < tol_infeas, or max_iters; the test runs every
        block_iters*check_every iterations.  Returns self.results()."""
        if block_iters < 1 or check_every < 1:
            raise ValueError("block_iters and check_every must be positive")
        tol_infeas = tol if tol_infeas is None else tol_infeas
        if int(self.iter_counter.item()) == 0 and int(self.nact.item()) == 0:
            self._compact(0)
        self.refresh_grid()
        graphs = {}
        warm = False
        done, since, nblk = 0, 0, 0
        while done < max_iters:
            n = min(block_iters, max_iters - done)
            if use_graph and n == block_iters:
                key = (self.grid, n)
                if key in graphs:
                    graphs[key].replay()
                elif not warm:
                    for _ in range(n):
                        self._iteration()
                    torch.cuda.synchronize()
                    warm = True
                else:
                    graphs[key] = self._capture(lambda: [self._iteration() for _ in range(n)])
                    graphs[key].replay()
            else:
                for _ in range(n):
                    self._iteration()
            done += n
            since += n
            nblk += 1
            if nblk % check_every == 0 or done >= max_iters:
                self._check(tol, tol_infeas, since)
                since = 0
                nrun = self.refresh_grid()
                stop = callback(self, done) if callback is not None else False
                if nrun == 0 or stop:
                    break
        self.iters_run = done
        return self.results()

    def results(self):
        chk = self.checks.cpu().numpy().reshape(self.B, CHECK_DOUBLES)
        iv = chk.view(np.int32)
        return [{"objp": c[0], "objd": c[1], "gap": c[2], "relgap": c[3], "infeasp": c[4], "infeasd": c[5],
                 "status": int(iv[b, 16]), "iter": int(iv[b, 17])} for b, c in enumerate(chk)]

    def solution(self):
        """Per instance (x, y) in the layout of PDHG.solve on lifting.sublabel_tv:
        x = (U (N, k), r (N, k)) and y = (p (N*k, 3), z (N*k, 2)), flat."""
        U, r = self.U.cpu().numpy().astype(np.float64), self.r.cpu().numpy().astype(np.float64)
        p, z = self.p.cpu().numpy().astype(np.float64), self.z.cpu().numpy().astype(np.float64)
        xs, ys = [], []
        for b, (H, W) in enumerate(self.shapes):
            N, k, ok = H * W, self.K[b], self.offks[b]
            Ub = U[ok:ok + N * k].reshape(k, N).T
            rb = r[ok:ok + N * k].reshape(k, N).T
            pb = p[2 * ok:2 * ok + 2 * N * k].reshape(k, 2, N).transpose(2, 0, 1)  # (N, k, 2)
            zb = z[2 * ok:2 * ok + 2 * N * k].reshape(k, 2, N).transpose(2, 0, 1)
            p3 = np.zeros((N, k, 3))
            p3[:, :, :2] = pb
            xs.append(np.concatenate([Ub.ravel(), rb.ravel()]))
            ys.append(np.concatenate([p3.ravel(), zb.ravel()]))
        return xs, ys

    def unlifted(self):
        """s = g_1 + h * sum_i U_i per instance, as (H, W) arrays."""
        U = self.U.cpu().numpy().astype(np.float64)
        out = []
        for b, (H, W) in enumerate(self.shapes):
            N, k, ok = H * W, self.K[b], self.offks[b]
            g = self.gam[b]
            out.append((g[0] + (g[1] - g[0]) * U[ok:ok + N * k].reshape(k, N).sum(axis=0)).reshape(H, W))
        return out
