"""Fused, batched PDHG for TV-regularised denoising of 2-D images.

One engine serves both the drop-in single-instance path (PDHG.solve on a
recognised TV-L2/TV-L1 model) and the batched solver over many instances of
uneven size.  Per iteration it launches one fused pass (two with passes=2)
and, with adaptive steps, one small per-instance rule kernel; every
`block_iters` iterations it evaluates the FP64 primal-dual gap, marks
converged instances and compacts
the list of running instances -- all on the device, so a whole block is one
hipGraph replay and the host only reads the number of running instances and
their tiles, which sizes the next launches (one block per running tile: the
hardware dispatcher then keeps the tiles in flight contiguous in memory,
which streams faster than a persistent grid whose blocks drift apart).

Model of instance b: min_x G_b(x) + tvw_b * sum_p |grad x(p)|_F with
G_b = 0.5*lam_b*|x - f_b|^2 (data="l2") or lam_b*|x - f_b|_1 (data="l1"),
the forward-difference gradient of the reference (diff.cu:64-71) and the
PDHG iteration of the reference (pdhg.py:124-187).
"""
import ctypes

import numpy as np
import torch

from opymize_amd import _hip

_DATA = {"l2": 0, "l1": 1}
# step rules: the reference's constant / adaptive (pdhg.py:145-183) and
# preconditioned (pdhg.py:189-209) steps, and the accelerated steps of
# Chambolle & Pock (2011, Algorithm 2) for the strongly convex TV-L2 instances
STEP_MODES = {"const": 0, "adaptive": 0, "precond": 1, "accel": 2}
ALIGN = 64  # elements; every instance starts on a 256-byte boundary (FP32)


class TVInst(ctypes.Structure):
    _fields_ = [("off", ctypes.c_int64), ("H", ctypes.c_int32), ("W", ctypes.c_int32),
                ("data", ctypes.c_int32), ("tile0", ctypes.c_int32), ("ntiles", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("lam", ctypes.c_double), ("tvw", ctypes.c_double)]


class TVCfg(ctypes.Structure):
    _fields_ = [("C", ctypes.c_int32), ("tw", ctypes.c_int32), ("rpt", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("h0", ctypes.c_double), ("h1", ctypes.c_double), ("w", ctypes.c_double * 4)]


class TVCheck(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("objp", "objd", "gap", "relgap", "infeasp", "infeasd", "objd_cert", "relgap_cert")] + \
               [("status", ctypes.c_int32), ("iters", ctypes.c_int32)]


class TVRestart(ctypes.Structure):
    _fields_ = [("ref", ctypes.c_double), ("best", ctypes.c_double), ("last", ctypes.c_int32),
                ("count", ctypes.c_int32)]


# restarts of the accelerated schedule: sufficient decay factor of the
# certified relative gap, and the fewest iterations between two restarts
RESTART_DEFAULT = (0.5, 96)

CHECK_DOUBLES = ctypes.sizeof(TVCheck) // 8
RESTART_DOUBLES = ctypes.sizeof(TVRestart) // 8
STATE_DOUBLES = _hip.ADAPT_STATE_DOUBLES
STREAM_MIN_UNITS = 4096
assert ctypes.sizeof(TVInst) == 48 and ctypes.sizeof(TVCfg) == 64 and ctypes.sizeof(TVCheck) == 72
assert ctypes.sizeof(TVRestart) == 24

_SIGS = {
    "opy_tv_primal": [ctypes.c_void_p] * 10 + [ctypes.c_int, ctypes.c_void_p],
    "opy_tv_dual": [ctypes.c_void_p] * 10 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_iter": [ctypes.c_void_p] * 11 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_multi": [ctypes.c_void_p] * 11 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_stream": [ctypes.c_void_p] * 11 + [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_gap": [ctypes.c_void_p] * 9 + [ctypes.c_int, ctypes.c_void_p],
    "opy_tv_reset": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4,
}
_PLAIN = {
    "opy_tv_adapt": [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
    "opy_tv_accel": [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p],
    "opy_tv_accel_n": [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_restart": [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_void_p],
    "opy_tv_check": [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    "opy_tv_compact": [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int,
                                                                                      ctypes.c_void_p],
    "opy_tv_stream_units": [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
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


def _stream_strips(W, sw):
    """Strips of the streaming kernel for an instance W columns wide (the
    kernel's stream_strips): ceil(W / sw) strips of sw = 64 - 2K exact
    columns, or, from K = 6 on when that takes fewer, the extended layout
    whose first and last strips need no halo beyond the image edge,
    64 + (n - 1) sw columns for n strips."""
    plain = -(-W // sw)
    ext = 1 if W <= 64 else -(-(W - 64) // sw) + 1
    return min(plain, ext) if sw <= 52 else plain


def tile_height(tw, rpt):
    return (256 // tw) * rpt


class TVBatch(object):
    """Batched fused TV PDHG.

    images: list of arrays (H, W) or (H, W, C) (all with the same C <= 4)
    lam:    data weight per instance (scalar or list)
    data:   "l2" / "l1" per instance (str or list)
    tvw:    TV weight per instance (scalar or list)
    passes: 1 = one kernel per iteration (reads x, f, y; writes x+, y+),
            2 = a primal and a dual pass (reads x, f, y, then x, x+, y)
    tile:   (tile width, rows per thread); with passes=1 the width is 64
    multi:  iterations per pass over HBM with the constant, preconditioned
            and accelerated step rules (passes=1): K in {2, 3, 4} runs K
            iterations per launch (on LDS-staged tiles of rows per thread 2,
            4, 6 or 8, or streamed, see multi_kernel), K in 5..8 on the
            streaming kernel only (C <= 2, no preconditioned steps), 1 one
            iteration per launch; adaptive steps always take one iteration
            per launch (their rule needs every iteration's residuals)
    multi_kernel: how K > 1 iterations per launch are split over the GPU:
            "stream" (one wavefront per strip of 64 - 2K columns by
            stream_rows rows, the K iterations pipelined down the rows), "tile"
            (LDS-staged tiles, K <= 4), or "auto": stream while the running
            instances give at least stream_min_units units (decided at each
            check), else tiles with min(K, 4) iterations per launch, and
            tiles for preconditioned steps; every choice gives the iterates
            of single launches bit for bit
    stream_rows: rows per unit of the streaming kernel
    """

    def __init__(self, images, lam, data="l2", tvw=1.0, precision="float", tile=(64, 4), h=(1.0, 1.0),
                 weights=None, grid=None, passes=1, multi=1, multi_kernel="auto", stream_rows=128):
        _hip.require_device()
        _bind()
        self.B = len(images)
        if self.B < 1:
            raise ValueError("need at least one instance")
        imgs = [np.asarray(im, dtype=np.float64) for im in images]
        imgs = [im[:, :, None] if im.ndim == 2 else im for im in imgs]
        C = imgs[0].shape[2]
        if any(im.shape[2] != C for im in imgs) or not 1 <= C <= 4:
            raise ValueError("all instances need the same channel count C in [1, 4]")
        self.C = C
        self.shapes = [im.shape[:2] for im in imgs]
        lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (self.B,))
        tvw = np.broadcast_to(np.asarray(tvw, dtype=np.float64), (self.B,))
        data = [data] * self.B if isinstance(data, str) else list(data)
        if np.any(lam <= 0) or np.any(tvw <= 0):
            raise ValueError("lam and tvw must be positive")
        self.lam, self.tvw, self.data = lam.copy(), tvw.copy(), data
        self.dtype = _hip.torch_dtype(precision)
        self.h = (float(h[0]), float(h[1]))
        self.weights = np.ones(C) if weights is None else np.asarray(weights, dtype=np.float64)
        # packed layout
        offs, o = [], 0
        for (H, W) in self.shapes:
            offs.append(o)
            o += ((H * W * C + ALIGN - 1) // ALIGN) * ALIGN
        self.offs = offs
        self.n_total = o
        fpack = np.zeros(o)
        for im, off in zip(imgs, offs):
            fpack[off:off + im.size] = im.ravel()
        self.f = _hip.to_device(fpack, self.dtype)
        self.x = [torch.zeros(o, dtype=self.dtype, device="cuda") for _ in range(2)]
        self.y = [torch.zeros(2 * o, dtype=self.dtype, device="cuda") for _ in range(2)]
        self._grid_override = grid
        if passes not in (1, 2):
            raise ValueError("passes must be 1 or 2")
        self.passes = passes
        if multi_kernel not in ("auto", "stream", "tile"):
            raise ValueError("multi_kernel must be 'auto', 'stream' or 'tile'")
        self.multi_kernel = multi_kernel
        self.unit_off = torch.zeros(self.B + 1, dtype=torch.int32, device="cuda")
        self.stream_rows = 128
        self.total_units, self.sgrid = 0, 0
        # "auto" streams while the running instances give at least this many
        # units (about four wavefronts per SIMD), else takes the tile kernel
        self.stream_min_units = STREAM_MIN_UNITS
        self._stream_now = True
        self.set_multi(multi)
        self.set_stream_rows(stream_rows)
        self.set_tile(*tile)
        self.graph = None
        self.cur = 0  # buffer holding the current iterate
        self._restart = None  # (beta, min_iters, tau0, sigma0) while a solve restarts its schedule

    # -- layout --------------------------------------------------------------
    def set_multi(self, k):
        if not 1 <= int(k) <= 8:
            raise ValueError("multi must be 1 to 8 iterations per launch")
        if int(k) > 4 and (self.multi_kernel == "tile" or self.C > 2):
            raise ValueError("multi > 4 runs on the streaming kernel only, for C <= 2")
        if int(k) > 1 and self.passes != 1:
            raise ValueError("multi > 1 needs passes=1")
        self.multi = int(k)
        self._stream_layout()

    def set_stream_rows(self, rows):
        """Rows per unit of the streaming multi-iteration kernel."""
        if not 1 <= int(rows) <= 1 << 20:
            raise ValueError("stream_rows must be in [1, 2^20]")
        self.stream_rows = int(rows)
        self._stream_layout()

    def _stream_layout(self):
        """Units of the streaming kernel over the whole batch (sizes its
        launches) and the running instances' unit prefix."""
        if self.multi < 2:
            self.total_units = 0
            return
        sw, rows = 64 - 2 * self.multi, self.stream_rows
        self.total_units = sum(_stream_strips(W, sw) * (-(-H // rows)) for (H, W) in self.shapes)
        if self._grid_override is None:
            self.sgrid = self._stream_grid(self.total_units)
        else:
            self.sgrid = int(self._grid_override)
        if hasattr(self, "act_off"):
            self._stream_units()

    def _stream_units(self):
        if self.multi > 1:
            P = _hip.ptr
            self._call("opy_tv_stream_units", P(self.insts), P(self.active_inst), P(self.nact), P(self.unit_off),
                       self.multi, self.stream_rows, _hip.stream())

    def _stream_grid(self, units):
        """Blocks (4 wavefronts, one unit each) for `units` running units,
        bucketed like _launch_grid and rounded up to a multiple of the 8
        XCDs; blocks past the running units return at once."""
        full = max(1, -(-self.total_units // 4))
        q = max(256, -(-full // 16))
        blocks = max(1, min(full, -(-max(-(-units // 4), 1) // q) * q))
        return -(-blocks // 8) * 8

    def _use_stream(self, cfg):
        """Whether the next K-iteration launches stream (else: LDS tiles)."""
        if self.multi_kernel == "tile" or self.multi < 2:
            return False
        if cfg.steps == STEP_MODES["precond"]:
            if self.multi_kernel == "stream" or self.multi > 4:
                raise ValueError("the streaming kernel has no preconditioned steps; use multi_kernel='tile' "
                                 "(at most 4 iterations per launch)")
            return False
        if not self._stream_fits():
            if self.multi_kernel == "stream" or self.multi > 4:
                raise ValueError("the streaming kernel addresses an instance's y with 32-bit byte offsets; the "
                                 "largest instance here needs more: use multi_kernel='tile' (at most 4 "
                                 "iterations per launch)")
            return False
        return self.multi_kernel == "stream" or self._stream_now

    def _stream_fits(self):
        """Every instance's y block (2 C values per pixel) is below 4 GiB, the
        reach of the streaming kernel's 32-bit offsets."""
        esize = 4 if self.dtype == torch.float32 else 8
        return max(H * W for (H, W) in self.shapes) * 2 * self.C * esize < (1 << 32)

    def _launch_k(self, cfg):
        """Iterations per K-iteration launch: multi, or at most 4 on the tile
        kernel (the same iterates, grouped differently)."""
        return self.multi if self._use_stream(cfg) else min(self.multi, 4)

    def set_tile(self, tw, rpt):
        if self.passes == 1:
            tw = 64
        if int(rpt) not in (1, 2, 4, 6, 8, 12, 16):
            raise ValueError("rows per thread %d not in {1, 2, 4, 6, 8, 12, 16}" % rpt)
        if self.C == 4 and int(rpt) == 16:
            raise ValueError("C=4 with 16 rows per thread exceeds the fused kernels' register budget; "
                             "use at most 12 rows per thread")
        th = tile_height(tw, rpt)
        self.tw, self.rpt = int(tw), int(rpt)
        insts = (TVInst * self.B)()
        t0 = 0
        self.ntiles = []
        for b, (H, W) in enumerate(self.shapes):
            nt = ((H + th - 1) // th) * ((W + tw - 1) // tw)
            I = insts[b]
            I.off, I.H, I.W = self.offs[b], H, W
            I.data = _DATA[self.data[b]]
            I.tile0, I.ntiles = t0, nt
            I.lam, I.tvw = float(self.lam[b]), float(self.tvw[b])
            self.ntiles.append(nt)
            t0 += nt
        self.total_tiles = t0
        raw = np.frombuffer(bytes(insts), dtype=np.uint8).copy()
        self.insts = torch.as_tensor(raw).cuda()
        self.part_res = torch.zeros(4 * t0, dtype=torch.float64, device="cuda")
        self.part_gap = torch.zeros(6 * t0, dtype=torch.float64, device="cuda")
        self.active_inst = torch.zeros(self.B, dtype=torch.int32, device="cuda")
        self.act_off = torch.zeros(self.B + 1, dtype=torch.int32, device="cuda")
        self.nact = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.iter_counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        # by default every launch gets one block per tile of the running
        # instances, rounded up to a bucket (see _launch_grid), which keeps the
        # tiles in flight contiguous; grid=0 makes the library size a
        # persistent launch to the blocks resident at once, grid=n > 0 fixes n
        if self._grid_override is None:
            self.grid = t0
        elif self._grid_override == 0:
            self.grid = 0
        else:
            self.grid = int(min(self._grid_override, t0))
        self.graph = None

    def _launch_grid(self, tiles):
        """Blocks per launch for `tiles` running tiles: one per tile, rounded
        up to a multiple of a sixteenth of the batch (at least 1024) so that a
        solve needs few distinct launch shapes (one captured graph each).
        Blocks past the running tiles return at once."""
        if self._grid_override is not None:
            return self.grid
        q = max(1024, -(-self.total_tiles // 16))
        return int(max(1, min(self.total_tiles, -(-max(tiles, 1) // q) * q)))

    def refresh_grid(self):
        """Read the running tile count (one device read) and size the
        following launches to it; returns the number of running instances."""
        nrun = int(self.nact.item())
        tiles = int(self.act_off[nrun].item()) if nrun > 0 else 0
        self.grid = self._launch_grid(tiles)
        if self.multi > 1:
            units = int(self.unit_off[nrun].item()) if nrun > 0 else 0
            if self._grid_override is None:
                self.sgrid = self._stream_grid(units)
            self._stream_now = units >= self.stream_min_units
        return nrun

    def _cfg(self, steps):
        c = TVCfg()
        c.C, c.tw, c.rpt = self.C, self.tw, self.rpt
        c.steps = STEP_MODES[steps]
        c.h0, c.h1 = self.h
        for k in range(4):
            c.w[k] = float(self.weights[k]) if k < self.C else 1.0
        return c

    # -- state -----------------------------------------------------------------
    def set_start(self, x0=None, y0=None):
        """Packed start values (device tensors or host arrays of the packed size), zeros otherwise."""
        self.cur = 0
        for t, v in ((self.x[0], x0), (self.y[0], y0)):
            if v is None:
                t.zero_()
            else:
                t.copy_(v.reshape(-1).to(self.dtype) if _hip.is_device_tensor(v)
                        else torch.as_tensor(np.asarray(v, dtype=_hip.np_dtype(self.dtype)).ravel()))

    def pack(self, arrays, field="x"):
        """Host list of per-instance arrays -> packed host vector."""
        mult = 1 if field == "x" else 2
        out = np.zeros(mult * self.n_total)
        for a, off in zip(arrays, self.offs):
            a = np.asarray(a, dtype=np.float64).ravel()
            out[mult * off:mult * off + a.size] = a
        return out

    def unpack(self, t, field="x"):
        mult = 1 if field == "x" else 2
        h = t.cpu().numpy().astype(np.float64)
        res = []
        for (H, W), off in zip(self.shapes, self.offs):
            n = H * W * self.C * mult
            res.append(h[mult * off:mult * off + n].copy())
        return res

    def _reset_state(self, steps, step_bound, step_factor, tol, tol_infeas, certify, statuses=None):
        st = np.zeros((self.B, STATE_DOUBLES))
        F = _hip.ADAPT_FIELD
        bnd = step_bound
        if steps == "adaptive":
            st[:, F["tau"]] = st[:, F["sigma"]] = np.sqrt(bnd)
            st[:, F["alpha"]] = 0.5
        else:
            sigma = np.sqrt(bnd / step_factor)
            st[:, F["sigma"]] = sigma
            st[:, F["tau"]] = bnd / sigma
        self.state = torch.as_tensor(st).cuda()
        chk = np.zeros((self.B, CHECK_DOUBLES))
        if statuses is not None:
            chk.view(np.int32)[:, 2 * 8] = statuses
        self.checks = torch.as_tensor(chk).cuda()
        rec = np.zeros((self.B, RESTART_DOUBLES))
        rec[:, :2] = np.inf  # ref, best: none yet; last = count = 0
        self.restarts = torch.as_tensor(rec).cuda()
        self.iter_counter.zero_()
        self._compact(0)

    # -- kernels ---------------------------------------------------------------
    def _sfx(self):
        return "_f32" if self.dtype == torch.float32 else "_f64"

    def _call(self, name, *args):
        lib = _hip.load()
        fn = getattr(lib, name)
        rc = fn(*args)
        if rc != 0:
            raise _hip.HipError("%s failed: %s" % (name, lib.opy_last_error().decode()))

    def _iteration(self, p, cfg, adapt, params):
        P = _hip.ptr
        s = _hip.stream()
        xo, xn = self.x[p], self.x[1 - p]
        yo, yn = self.y[p], self.y[1 - p]
        if self.passes == 1:
            self._call("opy_tv_iter" + self._sfx(), P(xo), P(xn), P(yo), P(yn), P(self.f), P(self.insts),
                       P(self.state), P(self.active_inst), P(self.act_off), P(self.nact), ctypes.byref(cfg),
                       int(adapt), P(self.part_res), self.grid, s)
        else:
            self._call("opy_tv_primal" + self._sfx(), P(xo), P(xn), P(yo), P(self.f), P(self.insts), P(self.state),
                       P(self.active_inst), P(self.act_off), P(self.nact), ctypes.byref(cfg), self.grid, s)
            self._call("opy_tv_dual" + self._sfx(), P(xo), P(xn), P(yo), P(yn), P(self.insts), P(self.state),
                       P(self.active_inst), P(self.act_off), P(self.nact), ctypes.byref(cfg), int(adapt),
                       P(self.part_res), self.grid, s)
        if adapt:
            self._call("opy_tv_adapt", P(self.insts), P(self.state), P(self.checks), P(self.part_res), self.B,
                       self.C, ctypes.byref(params), s)
        elif cfg.steps == STEP_MODES["accel"]:
            self._call("opy_tv_accel", P(self.insts), P(self.state), P(self.checks), self.B, s)

    def _gap_check(self, cfg, tol, tol_infeas, certify, iters_in_block):
        P = _hip.ptr
        s = _hip.stream()
        self._call("opy_tv_gap" + self._sfx(), P(self.x[self.cur]), P(self.y[self.cur]), P(self.f), P(self.insts),
                   P(self.active_inst), P(self.act_off), P(self.nact), ctypes.byref(cfg), P(self.part_gap),
                   self.grid, s)
        self._call("opy_tv_check", P(self.insts), P(self.checks), P(self.part_gap), self.B, float(tol),
                   float(tol_infeas), int(certify), P(self.iter_counter), int(iters_in_block), s)

    def _compact(self, iters_in_block):
        P = _hip.ptr
        self._call("opy_tv_compact", P(self.insts), P(self.checks), self.B, P(self.active_inst), P(self.act_off),
                   P(self.nact), P(self.iter_counter), int(iters_in_block), _hip.stream())
        self._stream_units()

    def _multi_launch(self, p, cfg, k):
        """k iterations in one launch from buffer p, then the k step updates."""
        P = _hip.ptr
        s = _hip.stream()
        if self._use_stream(cfg) and k == self.multi and k > 1:
            self._call("opy_tv_stream" + self._sfx(), P(self.x[p]), P(self.x[1 - p]), P(self.y[p]),
                       P(self.y[1 - p]), P(self.f), P(self.insts), P(self.state), P(self.active_inst),
                       P(self.unit_off), P(self.nact), ctypes.byref(cfg), int(k), self.stream_rows, self.sgrid, s)
        else:
            self._call("opy_tv_multi" + self._sfx(), P(self.x[p]), P(self.x[1 - p]), P(self.y[p]),
                       P(self.y[1 - p]), P(self.f), P(self.insts), P(self.state), P(self.active_inst),
                       P(self.act_off), P(self.nact), ctypes.byref(cfg), int(k), self.grid, s)
        if cfg.steps == STEP_MODES["accel"]:
            self._call("opy_tv_accel_n", P(self.insts), P(self.state), P(self.checks), self.B, int(k), s)

    def _iterations(self, n, cfg, adapt, params):
        """n iterations starting from buffer self.cur (updates self.cur)."""
        k = self._launch_k(cfg) if not adapt else 1
        while k > 1 and n >= k:
            self._multi_launch(self.cur, cfg, k)
            self.cur ^= 1
            n -= k
        for _ in range(n):
            self._iteration(self.cur, cfg, adapt, params)
            self.cur ^= 1

    def _check_step(self, cfg, tol, tol_infeas, certify, iters_since):
        self._gap_check(cfg, tol, tol_infeas, certify, iters_since)
        if self._restart is not None:
            beta, min_iters, tau0, sigma0 = self._restart
            self._call("opy_tv_restart", _hip.ptr(self.insts), _hip.ptr(self.state), _hip.ptr(self.checks),
                       _hip.ptr(self.restarts), self.B, float(beta), int(min_iters), float(tau0), float(sigma0),
                       _hip.stream())
        self._compact(iters_since)

    def _graph_block(self, cache, fn):
        """Run fn (a block of launches) through a graph captured once per
        (launch grid, buffer holding the iterate at the start).  The first
        block of a solve runs eagerly, which loads its kernels before any
        capture; a new key is captured (that records without running) and
        then replayed."""
        key = (self.grid, self.sgrid, self._stream_now, self.cur)
        if key in cache:
            g, end = cache[key]
            g.replay()
            self.cur = end
        elif not cache["warm"]:
            fn()
            torch.cuda.synchronize()
            cache["warm"] = True
        else:
            start = self.cur
            g = self._capture(fn)
            end, self.cur = self.cur, start
            cache[key] = (g, end)
            g.replay()
            self.cur = end

    def _to_buffer0(self):
        """Bring the current iterate to buffer 0 (the start and result buffer)."""
        if self.cur != 0:
            for a in (self.x, self.y):
                _hip.call("copy", self.dtype, _hip.ptr(a[1]), _hip.ptr(a[0]), a[0].numel(), _hip.stream())
            self.cur = 0

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
    def check_now(self, steps="adaptive", tol=1e-5, tol_infeas=1e-5, certify=True):
        """Gap of the current iterates of the running instances (no iteration)."""
        cfg = self._cfg(steps)
        self._gap_check(cfg, tol, tol_infeas, certify, 0)
        self._compact(0)
        return self.results()

    def _cold_restart(self, flags, steps, step_bound, step_factor):
        """Reset the flagged instances to x = y = 0 with fresh step state."""
        fl = torch.as_tensor(np.asarray(flags, dtype=np.int32)).cuda()
        self._call("opy_tv_reset" + self._sfx(), _hip.ptr(self.insts), _hip.ptr(fl), self.B, self.C,
                   _hip.ptr(self.x[self.cur]), _hip.ptr(self.y[self.cur]), _hip.ptr(self.checks), _hip.stream())
        st = self.state.cpu().numpy()
        F = _hip.ADAPT_FIELD
        for b in np.flatnonzero(flags):
            st[b, :] = 0.0
            if steps == "adaptive":
                st[b, F["tau"]] = st[b, F["sigma"]] = np.sqrt(step_bound)
                st[b, F["alpha"]] = 0.5
            else:
                st[b, F["sigma"]] = np.sqrt(step_bound / np.broadcast_to(step_factor, (self.B,))[b])
                st[b, F["tau"]] = step_bound / st[b, F["sigma"]]
        self.state.copy_(torch.as_tensor(st))
        # restart records (TVRestart: ref, best, last | count): none yet, last = now
        rec = self.restarts.cpu().numpy()
        iters = self.checks.cpu().numpy().view(np.int32)[:, 17]
        for b in np.flatnonzero(flags):
            rec[b, :2] = np.inf
            rec.view(np.int32)[b, 4] = iters[b]
        self.restarts.copy_(torch.as_tensor(rec))
        self._compact(0)

    def solve(self, steps="adaptive", step_bound=0.124, step_factor=1.0, tol=1e-5, tol_infeas=None,
              max_iters=20000, block_iters=50, check_every=1, certify=True, use_graph=True, eps_res=None,
              callback=None, check_start=False, keep_state=False, fallback=False, restart="auto",
              step_factor_l1=None):
        """Iterate until every instance is certified (|relgap| < tol and
        infeasibility < tol_infeas) or max_iters; the test runs every
        block_iters*check_every iterations.  Returns self.results().

        fallback=True (for warm starts): the gap of the start point is
        recorded, and an instance whose gap has not improved at the first
        test (or is not finite) restarts cold from x = y = 0.

        restart (accelerated steps): "auto" or True restarts the step
        schedule of a TV-L2 instance at its initial tau, sigma (keeping x,
        y) when, at a check at least min_iters iterations after its last
        restart, its certified relative gap has fallen below beta times the
        gap at that restart or has risen above the smallest gap since;
        (beta, min_iters) sets the two, RESTART_DEFAULT otherwise; False
        or None keeps the plain schedule.

        step_factor_l1 (accelerated steps): the initial split sigma =
        sqrt(step_bound / f), tau = step_bound / sigma of the TV-L1
        instances, which keep constant steps under this rule (step_factor
        otherwise).  Small values suit them: on the uneven 512-instance
        batch, 1/256 certifies 267 of 276 TV-L1 instances in a median of 48
        iterations, against 218 in a median of 864 at 4."""
        if steps not in STEP_MODES:
            raise ValueError("steps must be one of %s" % sorted(STEP_MODES))
        if restart not in ("auto", True, False, None):
            if len(restart) != 2 or not 0.0 <= float(restart[0]) < 1.0 or int(restart[1]) < 0:
                raise ValueError("restart must be 'auto', True, False, None or (beta in [0, 1), min_iters >= 0)")
        if step_factor_l1 is not None and not float(step_factor_l1) > 0.0:
            raise ValueError("step_factor_l1 must be positive")
        tol_infeas = tol if tol_infeas is None else tol_infeas
        if fallback:
            check_start = True
        self.fallbacks = []
        gap0 = None
        adapt = steps == "adaptive"
        cfg = self._cfg(steps)
        params = _hip.AdaptParams()
        e = tol if eps_res is None else eps_res
        params.eta, params.Delta = 0.95, 1.5
        params.eps_absp = params.eps_relp = params.eps_absd = params.eps_reld = e
        restart_factor = step_factor
        if steps == "accel" and step_factor_l1 is not None:
            step_factor = np.where(np.array([d == "l1" for d in self.data]), float(step_factor_l1),
                                   float(step_factor))
        if not keep_state:
            self._reset_state(steps, step_bound, step_factor, tol, tol_infeas, certify)
        self._restart = None
        if steps == "accel" and restart not in (False, None):
            beta, min_iters = RESTART_DEFAULT if restart in ("auto", True) else restart
            sigma0 = np.sqrt(step_bound / restart_factor)  # TV-L2 instances only restart
            self._restart = (float(beta), int(min_iters), step_bound / sigma0, sigma0)
        self.cur = 0
        if check_start:
            self._check_step(cfg, tol, tol_infeas, certify, 0)
            if fallback:
                r0 = self.results()
                gap0 = np.array([r["gap"] for r in r0])
                bad = np.array([r["status"] == 2 for r in r0])
                if bad.any():
                    self.fallbacks = list(np.flatnonzero(bad))
                    self._cold_restart(bad, steps, step_bound, step_factor)
            if self.refresh_grid() == 0:
                self.iters_run = 0
                return self.results()
        else:
            self.refresh_grid()
        graphs = use_graph
        # captured blocks per (launch grid, buffer holding the iterate at the start)
        g_it, g_ck = {"warm": False}, {"warm": False}
        done_iters, nblk = 0, 0
        since = 0
        while done_iters < max_iters:
            n = min(block_iters, max_iters - done_iters)
            if graphs and n == block_iters:
                self._graph_block(g_it, lambda: self._iterations(n, cfg, adapt, params))
            else:
                self._iterations(n, cfg, adapt, params)
            done_iters += n
            since += n
            nblk += 1
            if nblk % check_every == 0 or done_iters >= max_iters:
                if graphs and since == block_iters * check_every:
                    self._graph_block(g_ck, lambda: self._check_step(cfg, tol, tol_infeas, certify,
                                                                     block_iters * check_every))
                else:
                    self._check_step(cfg, tol, tol_infeas, certify, since)
                since = 0
                if gap0 is not None:
                    # first test after a warm start: cold restart where the gap did not improve
                    r1 = self.results()
                    flags = np.array([(r["status"] == 2) or (r["status"] == 0 and not r["gap"] < g0)
                                      for r, g0 in zip(r1, gap0)])
                    flags &= ~np.isin(np.arange(self.B), self.fallbacks)
                    if flags.any():
                        self.fallbacks += list(np.flatnonzero(flags))
                        self._cold_restart(flags, steps, step_bound, step_factor)
                    gap0 = None
                nrun = self.refresh_grid()
                stop = callback(self, done_iters) if callback is not None else False
                if nrun == 0 or stop:
                    break
        self.iters_run = done_iters
        self._to_buffer0()
        return self.results()

    def results(self):
        chk = self.checks.cpu().numpy()
        rv = self.restarts.cpu().numpy().view(np.int32)
        out = []
        iv = chk.view(np.int32)
        for b in range(self.B):
            c = chk[b]
            out.append({"objp": c[0], "objd": c[1], "gap": c[2], "relgap": c[3], "infeasp": c[4],
                        "infeasd": c[5], "objd_cert": c[6], "relgap_cert": c[7],
                        "status": int(iv[b, 16]), "iter": int(iv[b, 17]),
                        "restarts": int(rv[b, 5])})
        return out

    def steps_state(self):
        st = self.state.cpu().numpy()
        F = _hip.ADAPT_FIELD
        return [{k: float(st[b, F[k]]) for k in ("tau", "sigma", "alpha", "resp", "resd", "epsp", "epsd")}
                for b in range(self.B)]

    def solution(self):
        """Per-instance (x, y) host arrays of the current iterates (buffer 0)."""
        return self.unpack(self.x[self.cur], "x"), self.unpack(self.y[self.cur], "y")
