"""Primal-dual hybrid gradient solver on the GPU (reference: opymize/solvers/pdhg.py).

Solves min_x G(x) + F(Ax) with constant, adaptive (residual balancing) or
diagonally preconditioned steps.  Differences in mechanism (not in results)
from the reference:

* all iterates stay on the device; "advance" swaps buffers instead of copying;
* the overrelaxation and the dual step are one kernel;
* adaptive step sizes live in device memory and are updated by a kernel
  (same rule and constants), so no host round trip happens per iteration and
  blocks of iterations can be captured into a hipGraph;
* the primal-dual gap is reduced on the GPU in FP64 for both precisions.

Problems with a recognised structure (TV-L2 / TV-L1 over a forward-difference
gradient of a 2-D image) run through fused kernels instead, see
``opymize_amd.solvers.fused``.
"""
import logging

import numpy as np
import torch

from opymize_amd import _hip
from opymize_amd.linear import normest
from opymize_amd.tools import truncate
from opymize_amd.tools.util import GracefulInterruptHandler


def _device_step_ok(F):
    """True if F.prox accepts a step read from device memory."""
    from opymize_amd.functionals import SplitSum
    if isinstance(F, SplitSum):
        return all(_device_step_ok(c) for c in F.fcts)
    return bool(getattr(F, "device_step", False))


def _as_host(v):
    return v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)


def fused_tv_model(g, f, A):
    """The TV-L2 / TV-L1 denoising structure the fused kernels implement, or None:
    G = SSD(data, vol) with equal volumes, no mask and no shift, or L1Data with
    equal weights; F = L1Norms(N, M, lbd) with the Frobenius norm and
    M1*M2 = 2*C; A = GradientOp(2-D, C <= 4, scheme="forward")."""
    from opymize_amd.functionals import SSD, L1Data, L1Norms
    from opymize_amd.linear import GradientOp
    if not isinstance(A, GradientOp) or A.scheme != "forward" or len(A.imagedims) != 2 or A.nchannels > 4:
        return None
    C = A.nchannels
    N = int(np.prod(A.imagedims))
    if not isinstance(f, L1Norms) or f.matrixnorm != "frobenius" or f.N != N or f.M[0] * f.M[1] != 2 * C:
        return None
    if isinstance(g, SSD):
        if not g.mask.all() or g.shift != 0 or not np.all(g.vol == g.vol[0]):
            return None
        data, lam = "l2", float(g.vol[0])
    elif isinstance(g, L1Data):
        if not np.all(g.lbd == g.lbd[0]):
            return None
        data, lam = "l1", float(g.lbd[0])
    else:
        return None
    if g.f.shape != (N, C):
        return None
    H, W = A.imagedims
    return {"data": data, "lam": lam, "tvw": f.lbd, "image": g.f.reshape(H, W, C), "h": A.imageh, "w": A.weights}


class GapEvaluator(object):
    """P = G(x) + F(Ax) and D = -G*(-A^T y) - F*(y) in FP64 on the device
    with one device-to-host copy (reference: pdhg.py:23-44)."""

    def __init__(self, g, f):
        self.g, self.f = g, f
        self.fns = [g, f, g.conj, f.conj]
        self.slots = [F.nslots() for F in self.fns]
        self.out = None
        self.neg = None

    def __call__(self, x, Ax, ATy, y):
        if self.out is None or self.out.device != x.device:
            self.out = torch.zeros(3 * sum(self.slots), dtype=torch.float64, device=x.device)
        if self.neg is None or self.neg.dtype != ATy.dtype or self.neg.numel() != ATy.numel():
            self.neg = torch.empty_like(ATy)
        _hip.call("scale", ATy.dtype, _hip.ptr(ATy), _hip.ptr(self.neg), ATy.numel(), None, -1.0, 0,
                  _hip.stream())
        args = [x, Ax, self.neg, y]
        o = 0
        for F, k, a in zip(self.fns, self.slots, args):
            F._gpu_eval(a, self.out[3 * o:3 * (o + k)])
            o += k
        v = self.out.cpu().numpy()
        res, o = [], 0
        for F, k in zip(self.fns, self.slots):
            res.append(F._gpu_finish(v[3 * o:3 * (o + k)]))
            o += k
        (gv, gi), (fv, fi), (gcv, gci), (fcv, fci) = res
        obj_p, infeas_p = gv + fv, gi + fi
        obj_d, infeas_d = -gcv - fcv, gci + fci
        return {
            "objp": obj_p, "objd": obj_d, "infeasp": infeas_p, "infeasd": infeas_d,
            "gap": obj_p - obj_d,
            "relgap": (obj_p - obj_d) / max(np.spacing(1), abs(obj_d)),
        }


class PDHG(object):
    """PDHG(g, f, A).solve(...) with the reference's signature and info dict
    (reference: pdhg.py:11-340)."""

    def __init__(self, g, f, A):
        assert A.x.size == g.x.size
        assert A.y.size == f.x.size
        self.g = g
        self.f = f
        self.linop = A
        self.itervars = {}
        self.constvars = {}
        self.use_gpu = True
        self.info = {}
        self._dev = {}
        self._fused = None

    # -- objective -----------------------------------------------------------
    def obj_primal(self, x, Ax):
        obj, infeas = self.g(x)
        obj2, infeas2 = self.f(Ax)
        return obj + obj2, infeas + infeas2

    def obj_dual(self, ATy, y):
        if _hip.is_device_tensor(ATy):
            neg = -1.0 * ATy  # host-side convenience path, not used inside solve
        else:
            neg = -np.asarray(ATy)
        obj, infeas = self.g.conj(neg)
        obj2, infeas2 = self.f.conj(y)
        return -obj - obj2, infeas + infeas2

    def pd_gap(self):
        d = self._dev
        return self._gap(d["xk"], d["ygradk"], d["xgradk"], d["yk"])

    # -- steps -------------------------------------------------------------
    def prepare_stepsizes(self, step_bound, step_factor, steps, precond_weight=1.0):
        i, c, d = self.itervars, self.constvars, self._dev
        dtype = d["xk"].dtype
        if steps == "precond":
            c["precond"] = True
            xtau = torch.zeros_like(d["xk"])
            ysigma = torch.zeros_like(d["yk"])
            logging.info("Determining diagonal preconditioners...")
            self.linop.rowwise_lp(ysigma)
            self.linop.adjoint.rowwise_lp(xtau)
            eps = float(np.spacing(1))
            _hip.call("recip_gt", dtype, _hip.ptr(ysigma), ysigma.numel(), eps, _hip.stream())
            _hip.call("recip_gt", dtype, _hip.ptr(xtau), xtau.numel(), eps, _hip.stream())
            # A prox that projects onto a set coupling a block of entries (an epigraph, a
            # simplex) is the exact prox only for a step that is constant on the block: the
            # reference's per-element steps give such a PDHG a fixed point off the solution.
            # Each block takes its smallest positive step, which keeps the convergence bound.
            for vec, fct in ((ysigma, self.f.conj), (xtau, self.g)):
                for off, nb, w in fct.precond_blocks():
                    _hip.call("block_min", dtype, _hip.addr(vec, off), nb, w, _hip.stream())
            if precond_weight != 1.0:
                # primal weight w (no reference counterpart): steps w*T and Sigma/w leave
                # |Sigma^1/2 K T^1/2| unchanged, so convergence holds for every w > 0
                if not precond_weight > 0:
                    raise ValueError("precond_weight must be positive")
                _hip.call("scale", dtype, _hip.ptr(xtau), _hip.ptr(xtau), xtau.numel(), None, float(precond_weight),
                          0, _hip.stream())
                _hip.call("scale", dtype, _hip.ptr(ysigma), _hip.ptr(ysigma), ysigma.numel(), None,
                          1.0 / float(precond_weight), 0, _hip.stream())
            d["xtau"], d["ysigma"] = xtau, ysigma
            self.gprox = self.g.prox(xtau)
            self.fconjprox = self.f.conj.prox(ysigma)
            return
        bnd = step_bound
        if step_bound is None:
            logging.info("Estimating optimal step bound...")
            op_norm, itn = normest(self.linop)
            bnd = truncate(1.0 / op_norm ** 2, 3)
        if steps == "adaptive":
            c["adaptive"] = True
            c["eta"] = 0.95
            c["Delta"] = 1.50
            st = torch.zeros(_hip.ADAPT_STATE_DOUBLES, dtype=torch.float64)
            st[_hip.ADAPT_FIELD["tau"]] = np.sqrt(bnd)
            st[_hip.ADAPT_FIELD["sigma"]] = np.sqrt(bnd)
            st[_hip.ADAPT_FIELD["alpha"]] = 0.5
            d["adapt"] = st.to("cuda")
            i["alphak"] = 0.5
            i["sigmak"] = i["tauk"] = np.sqrt(bnd)
            self._adapt_dev_steps = _device_step_ok(self.g) and _device_step_ok(self.f.conj)
            if self._adapt_dev_steps:
                self.gprox = self.g.prox((d["adapt"], _hip.ADAPT_FIELD["tau"]))
                self.fconjprox = self.f.conj.prox((d["adapt"], _hip.ADAPT_FIELD["sigma"]))
            else:
                self.gprox = self.g.prox(i["tauk"])
                self.fconjprox = self.f.conj.prox(i["sigmak"])
            logging.info("Adaptive steps: %f" % (bnd,))
        else:
            fact = step_factor
            c["sigma"] = np.sqrt(bnd / fact)
            c["tau"] = bnd / c["sigma"]
            self.gprox = self.g.prox(c["tau"])
            self.fconjprox = self.f.conj.prox(c["sigma"])
            logging.info("Constant steps: %f (%f | %f)" % (bnd, c["sigma"], c["tau"]))

    def _steps(self):
        c, d = self.constvars, self._dev
        if "precond" in c:
            return d["xtau"], d["ysigma"]
        if "adaptive" in c:
            st = d["adapt"]
            return (st, _hip.ADAPT_FIELD["tau"]), (st, _hip.ADAPT_FIELD["sigma"])
        return c["tau"], c["sigma"]

    # -- one iteration (reference pdhg.py:124-187) -------------------------
    def iteration_step(self):
        d, c = self._dev, self.constvars
        xk, xkp1, xgradk, xgradkp1 = d["xk"], d["xkp1"], d["xgradk"], d["xgradkp1"]
        yk, ykp1, ygradk, ygradkp1 = d["yk"], d["ykp1"], d["ygradk"], d["ygradkp1"]
        dtype = xk.dtype
        theta = c["theta"]
        tau, sigma = self._steps()
        stream = _hip.stream()

        # primal: x+ = prox_tauG(x - tau*A^T y); A x+
        st = _hip.make_step(tau, dtype)
        _hip.call("pdhg_take_step", dtype, _hip.ptr(xkp1), _hip.ptr(xk), xk.numel(), -1, _hip.ctypes.byref(st),
                  _hip.ptr(xgradk), stream)
        self.gprox(xkp1)
        self.linop(xkp1, ygradkp1)
        # dual: y+ = prox_sigmaF*(y + sigma*((1+theta) A x+ - theta A x)); A^T y+
        ss = _hip.make_step(sigma, dtype)
        _hip.call("pdhg_dual_step", dtype, _hip.ptr(ykp1), _hip.ptr(yk), yk.numel(), float(theta),
                  _hip.ctypes.byref(ss), _hip.ptr(ygradkp1), _hip.ptr(ygradk), stream)
        self.fconjprox(ykp1)
        self.linop.adjoint(ykp1, xgradkp1)

        if "adaptive" in c:
            adapt = d["adapt"]
            work = _hip.Workspace.get()
            sums = _hip.addr(adapt, _hip.ADAPT_FIELD["sums"])
            sums2 = _hip.addr(adapt, _hip.ADAPT_FIELD["sums"] + 2)
            st = _hip.make_step(tau, dtype)
            _hip.call("pdhg_residual2", dtype, _hip.ctypes.byref(st), _hip.ptr(xk), _hip.ptr(xkp1), 1.0,
                      _hip.ptr(xgradk), 1.0, 0.0, _hip.ptr(xgradkp1), xk.numel(), _hip.ptr(work), sums, stream)
            ss = _hip.make_step(sigma, dtype)
            _hip.call("pdhg_residual2", dtype, _hip.ctypes.byref(ss), _hip.ptr(yk), _hip.ptr(ykp1), float(theta),
                      _hip.ptr(ygradk), float(theta), float(theta) + 1.0, _hip.ptr(ygradkp1), yk.numel(),
                      _hip.ptr(work), sums2, stream)
            _hip.call_plain("opy_pdhg_adapt", _hip.ptr(adapt), _hip.ctypes.byref(self._adapt_params), stream)
            if not self._adapt_dev_steps:
                self._sync_host_steps()

        # advance: swap roles instead of copying
        d["xk"], d["xkp1"] = xkp1, xk
        d["xgradk"], d["xgradkp1"] = xgradkp1, xgradk
        d["yk"], d["ykp1"] = ykp1, yk
        d["ygradk"], d["ygradkp1"] = ygradkp1, ygradk

    def _sync_host_steps(self):
        """Functionals whose prox cannot read a device step: rebuild the prox
        maps on the host when the device rule changed the steps."""
        i = self.itervars
        st = self._dev["adapt"].cpu().numpy()
        n = int(st[_hip.ADAPT_FIELD["nchanges"]:_hip.ADAPT_FIELD["nchanges"] + 1].view(np.int32)[0])
        if n != self._host_nchanges:
            self._host_nchanges = n
            i["tauk"] = float(st[_hip.ADAPT_FIELD["tau"]])
            i["sigmak"] = float(st[_hip.ADAPT_FIELD["sigma"]])
            i["alphak"] = float(st[_hip.ADAPT_FIELD["alpha"]])
            self.gprox = self.g.prox(i["tauk"])
            self.fconjprox = self.f.conj.prox(i["sigmak"])

    def _read_adapt_info(self):
        st = self._dev["adapt"].cpu().numpy()
        F = _hip.ADAPT_FIELD
        for k in ("resp", "resd", "epsp", "epsd"):
            self.info[k] = float(st[F[k]])
        self.itervars["tauk"] = float(st[F["tau"]])
        self.itervars["sigmak"] = float(st[F["sigma"]])
        self.itervars["alphak"] = float(st[F["alpha"]])

    # -- graphs ------------------------------------------------------------
    def _build_graph(self):
        """Capture two iterations (the buffer roles are back in place after
        two swaps) into one hipGraph."""
        d = self._dev
        names = ["xk", "xkp1", "xgradk", "xgradkp1", "yk", "ykp1", "ygradk", "ygradkp1"]
        before = {k: d[k] for k in names}
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.iteration_step()
                self.iteration_step()
        torch.cuda.current_stream().wait_stream(s)
        assert all(d[k] is before[k] for k in names)
        return g

    # -- solve -------------------------------------------------------------
    def solve(self, continue_at=None, precision="double", steps="const", step_bound=None, step_factor=1.0,
              term_pd_gap=1e-5, term_pd_res=None, term_maxiter=int(5e4), granularity=5000, use_gpu=True,
              cbfun=None, use_graph=True, fused="auto", precond_weight=1.0):
        if not use_gpu:
            raise ValueError("opymize_amd implements the GPU path of PDHG (use_gpu=True)")
        _hip.require_device()
        lift = getattr(self.linop, "lifting_model", None)
        if fused and lift is not None and steps == "precond" and term_pd_res is None and term_maxiter > 0 \
                and continue_at is None:
            return self._solve_fused_lift(lift, precision, term_pd_gap, term_maxiter, granularity, cbfun, use_graph,
                                          precond_weight)
        if fused and term_pd_res is None and term_maxiter > 0:
            model = fused_tv_model(self.g, self.f, self.linop)
            if model is not None:
                return self._solve_fused(model, continue_at, precision, steps, step_bound, step_factor,
                                         term_pd_gap, term_maxiter, granularity, cbfun, use_graph)
            if fused is True:
                raise ValueError("fused=True but the model is not a 2-D TV-L2/TV-L1 problem")
        i, c, d = self.itervars, self.constvars, self._dev
        dtype = _hip.torch_dtype(precision)
        npdt = _hip.np_dtype(dtype)
        xk = torch.zeros(self.g.x.size, dtype=dtype, device="cuda")
        yk = torch.zeros(self.f.x.size, dtype=dtype, device="cuda")
        if continue_at is not None:
            x0, y0 = continue_at
            xk.copy_(torch.as_tensor(np.asarray(x0, dtype=npdt).ravel()) if not _hip.is_device_tensor(x0)
                     else x0.reshape(-1).to(dtype))
            yk.copy_(torch.as_tensor(np.asarray(y0, dtype=npdt).ravel()) if not _hip.is_device_tensor(y0)
                     else y0.reshape(-1).to(dtype))
        d["xk"], d["yk"] = xk, yk
        i["xk"], i["yk"] = xk.cpu().numpy(), yk.cpu().numpy()
        if term_maxiter == 0:
            self.info["iter"] = 0
            return self.info

        for s in ["xkp1", "xgradk", "xgradkp1"]:
            d[s] = xk.clone()
        for s in ["ykp1", "ygradk", "ygradkp1"]:
            d[s] = yk.clone()
        mem = (xk.numel() * xk.element_size() * 4 + yk.numel() * yk.element_size() * 4) // (1024 * 1024)
        logging.info("%d/%d (primal/dual) variables, %d MB of device memory" % (xk.numel(), yk.numel(), mem))

        self.linop(d["xk"], d["ygradk"])
        self.linop.adjoint(d["yk"], d["xgradk"])

        pd_res_mode = term_pd_res is not None
        if pd_res_mode:
            logging.info("Assuming adaptive step sizes (for pd_res_mode)")
            steps = "adaptive"
            if not isinstance(term_pd_res, tuple):
                term_pd_res = (term_pd_res,) * 4
            else:
                assert len(term_pd_res) == 4
            self.info["epspd"] = term_pd_res
        else:
            if not isinstance(term_pd_gap, tuple):
                term_pd_gap = (term_pd_gap, term_pd_gap)
            if steps == "adaptive":
                self.info["epspd"] = (term_pd_gap[0],) * 4

        c["theta"] = 1.0
        self.prepare_stepsizes(step_bound, step_factor, steps, precond_weight)
        if "adaptive" in c:
            p = _hip.AdaptParams()
            p.eta, p.Delta = c["eta"], c["Delta"]
            p.sqrt_n, p.sqrt_m = np.sqrt(xk.numel()), np.sqrt(yk.numel())
            p.eps_absp, p.eps_relp, p.eps_absd, p.eps_reld = self.info["epspd"]
            self._adapt_params = p
            self._host_nchanges = 0
        self._gap = GapEvaluator(self.g, self.f)

        graph = None
        can_graph = use_graph and not ("adaptive" in c and not self._adapt_dev_steps)

        logging.info("Solving (steps<%d)..." % term_maxiter)
        self.log_header(pd_res_mode)
        with GracefulInterruptHandler() as interrupt_hdl:
            _iter = 0
            while _iter < term_maxiter:
                # iterations up to the next check (or the end)
                next_check = min(term_maxiter, (_iter // granularity + 1) * granularity)
                todo = next_check - _iter
                if can_graph and todo >= 4:
                    if graph is None:
                        # eager warm-up pair: uploads constants before capture
                        self.iteration_step()
                        self.iteration_step()
                        todo -= 2
                        _iter += 2
                        graph = self._build_graph()
                    while todo >= 2 and not interrupt_hdl.interrupted:
                        graph.replay()
                        todo -= 2
                        _iter += 2
                while todo > 0 and not interrupt_hdl.interrupted:
                    self.iteration_step()
                    todo -= 1
                    _iter += 1
                check_step = (_iter % granularity == 0)
                if interrupt_hdl.interrupted or check_step:
                    if interrupt_hdl.interrupted:
                        print("Interrupt (SIGINT) at iter=%d" % _iter)
                    self.info["iter"] = _iter
                    if "adaptive" in c:
                        self._read_adapt_info()
                    i["xk"] = d["xk"].cpu().numpy()
                    i["yk"] = d["yk"].cpu().numpy()
                    if pd_res_mode:
                        test_term = self.info["resp"] < self.info["epsp"] and self.info["resd"] < self.info["epsd"]
                    else:
                        self.info.update(self.pd_gap())
                        test_err = np.abs(self.info["relgap"]) < term_pd_gap[0]
                        test_infeas = max(self.info["infeasp"], self.info["infeasd"]) < term_pd_gap[1]
                        test_term = test_err and test_infeas
                    self.log_iter(pd_res_mode)
                    if cbfun is not None:
                        cbfun(self.state, self.info)
                    if test_term or interrupt_hdl.interrupted:
                        break
        torch.cuda.synchronize()
        return self.info

    def _solve_fused(self, model, continue_at, precision, steps, step_bound, step_factor, term_pd_gap,
                     term_maxiter, granularity, cbfun, use_graph):
        """Same loop semantics as solve(), through the fused TV kernels."""
        from opymize_amd.solvers.tv import TVBatch
        from opymize_amd.solvers import autotune
        if not isinstance(term_pd_gap, tuple):
            term_pd_gap = (term_pd_gap, term_pd_gap)
        if steps not in ("precond", "adaptive"):
            steps_mode = "const"
        else:
            steps_mode = steps
        bnd = step_bound
        if steps_mode != "precond" and bnd is None:
            logging.info("Estimating optimal step bound...")
            op_norm, _ = normest(self.linop)
            bnd = truncate(1.0 / op_norm ** 2, 3)
        H, W, C = model["image"].shape
        tile = autotune.default_tile(H, W, C)
        eng = TVBatch([model["image"]], model["lam"], model["data"], model["tvw"], precision=precision,
                      tile=tile, h=model["h"], weights=model["w"])
        if continue_at is not None:
            x0, y0 = continue_at
            eng.set_start(eng.pack([_as_host(x0)], "x"), eng.pack([_as_host(y0)], "y"))
        if granularity % 50 == 0:
            block, every = 50, granularity // 50
        else:
            block, every = granularity, 1
        self.constvars = {"theta": 1.0, "fused": True}
        if steps_mode == "adaptive":
            self.constvars.update(adaptive=True, eta=0.95, Delta=1.5)
        elif steps_mode == "precond":
            self.constvars["precond"] = True
        logging.info("Fused TV kernels (%s data term, %s steps, tile %dx%d)"
                     % (model["data"], steps_mode, eng.tw, eng.rpt))
        self.log_header(False)
        self._fused = eng
        with GracefulInterruptHandler() as interrupt_hdl:
            def cb(e, it):
                r = e.results()[0]
                self.info.update({k: r[k] for k in ("objp", "objd", "infeasp", "infeasd", "gap", "relgap")})
                self.info["iter"] = it
                if steps_mode == "adaptive":
                    s = e.steps_state()[0]
                    for k in ("resp", "resd", "epsp", "epsd"):
                        self.info[k] = s[k]
                    self.itervars.update(tauk=s["tau"], sigmak=s["sigma"], alphak=s["alpha"])
                xs, ys = e.solution()
                self.itervars["xk"], self.itervars["yk"] = xs[0], ys[0]
                self.log_iter(False)
                if cbfun is not None:
                    cbfun(self.state, self.info)
                return interrupt_hdl.interrupted
            eng.solve(steps=steps_mode, step_bound=bnd if bnd is not None else 1.0, step_factor=step_factor,
                      tol=term_pd_gap[0], tol_infeas=term_pd_gap[1], max_iters=term_maxiter, block_iters=block,
                      check_every=every, certify=False, use_graph=use_graph, callback=cb)
        torch.cuda.synchronize()
        return self.info

    def _solve_fused_lift(self, model, precision, term_pd_gap, term_maxiter, granularity, cbfun, use_graph,
                          precond_weight=1.0):
        """The multi-label lifting model of opymize_amd.lifting through the fused
        kernels (same iterates as the generic chain with steps="precond")."""
        from opymize_amd.solvers.lift import LiftBatch
        if not isinstance(term_pd_gap, tuple):
            term_pd_gap = (term_pd_gap, term_pd_gap)
        eng = LiftBatch([model["f"]], model["L"], model["lam"], lo=model["lo"], hi=model["hi"],
                        precision=precision, omega=precond_weight)
        block = 100 if granularity % 100 == 0 else granularity
        self.constvars = {"theta": 1.0, "precond": True, "fused": True}
        logging.info("Fused lifting kernels (%d labels, preconditioned steps)" % model["L"])
        self.log_header(False)
        self._fused = eng
        with GracefulInterruptHandler() as interrupt_hdl:
            def cb(e, it):
                r = e.results()[0]
                self.info.update({k: r[k] for k in ("objp", "objd", "infeasp", "infeasd", "gap", "relgap")})
                self.info["iter"] = it
                xs, ys = e.solution()
                self.itervars["xk"], self.itervars["yk"] = xs[0], ys[0]
                self.log_iter(False)
                if cbfun is not None:
                    cbfun(self.state, self.info)
                return interrupt_hdl.interrupted
            eng.solve(max_iters=term_maxiter, block_iters=block, check_every=granularity // block,
                      tol=term_pd_gap[0], tol_infeas=term_pd_gap[1], use_graph=use_graph, callback=cb)
        torch.cuda.synchronize()
        return self.info

    def log_header(self, pd_res_mode):
        if pd_res_mode:
            logging.info("#{:>7s}: {:>13s} ({:>9s}) {:>13s} ({:>9s})".format(
                "iter", "primal resid.", "tolerance", "dual resid.", "tolerance"))
        else:
            logging.info("#{:>7s}: {:>9s} ({:>8s}) {:>9s} ({:>8s}) {:>9s} ({:>9s})".format(
                "iter", "primal", "infeas", "dual", "infeas", "pd-gap", "relative"))

    def log_iter(self, pd_res_mode):
        i = self.info
        if pd_res_mode:
            logging.info("#{:7d}: {:13.4g} ({:9.3e}) {:13.4g} ({:9.3e})".format(
                i["iter"], i["resp"], i["epsp"], i["resd"], i["epsd"]))
        else:
            logging.info("#{:7d}: {:9.4g} ({:8.2e}) {:9.4g} ({:8.2e}) {:9.3g} ({:9.2e})".format(
                i["iter"], i["objp"], i["infeasp"], i["objd"], i["infeasd"], i["gap"], i["relgap"]))

    @property
    def state(self):
        return (self.itervars["xk"], self.itervars["yk"])
