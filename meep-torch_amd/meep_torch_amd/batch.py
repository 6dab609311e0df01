"""Batched 2D TM design variants of uneven size on one GPU.

BatchedTM advances B independent TM problems (each its own grid size, eps,
CPML, source and DFT monitors) together: every block of steps is ONE
persistent launch pair over the (variant, tile) work list of all variants
(mt_tm2d_batch_run), so small and large grids share the chip.  All buffers are
allocated once; `capture()` records the whole forward run (checkpoint copies,
stepping, spectra, per-chunk folds) in a HIP graph that every later forward
replays, the pattern of an inverse-design loop where only eps changes.

The gradient with respect to each variant's eps runs the batched reverse-time
adjoint (mt_tm2d_batch_adjoint) chunk by chunk over the checkpoints, with the
chunk recomputed on a per-variant Ez tape; `BatchedTM.apply(eps_list)` wraps
forward + adjoint in torch.autograd.Function.

Monitor accumulators follow run_tm's semantics with checkpoint_every = K: each
K-step chunk is summed separately and chunk sums are added in order
(meep_torch/core/run.py:56-63), so a variant's spectra equal run_tm's.
"""
import ctypes

import torch

from . import _engine as E
from .core import engine2d
from .core.stepper_tm import FIELDS

_ORDER = engine2d.TM_ORDER


class BatchItem(ctypes.Structure):
    _fields_ = [("prob", E.TM2D), ("src", E.Source2D), ("coef_tab", E.c_void_p),
                ("mons", ctypes.POINTER(E.Monitor)), ("n_mons", E.c_int), ("pad_", E.c_int),
                ("stage", E.c_void_p), ("stage_bytes", E.c_ll), ("hcoef_tab", E.c_void_p)]


def _bind(L):
    if getattr(L, "_batch_bound", False):
        return
    P = ctypes.POINTER
    L.mt_tm2d_batch_workspace.restype = E.c_ll
    L.mt_tm2d_batch_workspace.argtypes = [E.c_int, E.c_int]
    L.mt_tm2d_batch_create.restype = E.c_void_p
    L.mt_tm2d_batch_create.argtypes = [P(BatchItem), E.c_int, E.c_int, E.c_int, E.c_void_p, E.c_ll, E.c_void_p]
    L.mt_tm2d_batch_run.restype = E.c_int
    L.mt_tm2d_batch_run.argtypes = [E.c_void_p, E.c_int, E.c_int, E.c_void_p]
    L.mt_tm2d_batch_adjoint.restype = E.c_int
    L.mt_tm2d_batch_adjoint.argtypes = [E.c_void_p, P(E.TM2DAdj), P(E.c_void_p), P(E.Monitor),
                                        P(E.c_int), E.c_int, E.c_int, E.c_void_p]
    L.mt_tm2d_batch_adjoint_prepare.restype = E.c_int
    L.mt_tm2d_batch_adjoint_prepare.argtypes = [E.c_void_p, P(E.TM2DAdj), P(E.c_void_p), P(E.Monitor), P(E.c_int),
                                                E.c_void_p]
    L.mt_tm2d_batch_adjoint_sweep.restype = E.c_int
    L.mt_tm2d_batch_adjoint_sweep.argtypes = [E.c_void_p, E.c_int, E.c_int, E.c_void_p]
    L.mt_tm2d_batch_adjoint_pingpong.restype = E.c_int
    L.mt_tm2d_batch_adjoint_pingpong.argtypes = [E.c_void_p, P(E.TM2DAdj), P(E.Monitor), P(E.c_int), P(E.c_int),
                                                  E.c_void_p]
    L.mt_batch_destroy.restype = None
    L.mt_batch_destroy.argtypes = [E.c_void_p]
    L._batch_bound = True


def _has_media(c):
    return bool(c.c1.shape[0] or c.has_kerr or c.has_chi2 or c.d1 is not None or c.h1x is not None
                or c.mc1x is not None)


class BatchedTM:
    """B TM problems stepped together.  problems: list of (coeffs, source,
    monitors) with coeffs from build_tm_coeffs (eps_inf is replaced per call).
    Every variant runs n_steps with checkpoints every K steps.  With
    adjoint=False no checkpoints or Ez tapes are kept (forward runs only;
    spectra are still folded per K-step chunk).

    Variants with dispersive, nonlinear or conductive media (ADE poles,
    chi2 / chi3, sigma_d, sigma_b / mu, magnetic poles) run one after the
    other through the single-problem path (its two-pass media kernels and,
    for gradients, its hand-written adjoint with checkpoints every K steps),
    with the same forward / gradient / apply interface; a batch holding
    such a variant takes that path for every variant.  Graph capture is for
    the batched launch only."""

    def __init__(self, problems, n_steps, checkpoint_every=200, variant=None, stage_steps=64, adjoint=True,
                 fused_adjoint=True):
        self.L = E.lib()
        # reverse steps as one fused pass over tiles with two cotangent sets
        # (False: the three transpose passes; the gradient is the same bit for bit)
        self.fused_adjoint = bool(fused_adjoint)
        _bind(self.L)
        self.B = len(problems)
        self.n_steps = int(n_steps)
        self.K = int(checkpoint_every) if checkpoint_every > 0 else self.n_steps
        c0 = problems[0][0]
        self.dtype = c0.eps_inf.dtype
        self.device = c0.eps_inf.device
        E.require_device(c0.eps_inf)
        self.code = E.dtype_code(self.dtype)
        if variant is None:
            # FP32: 32 x 64 tiles, 4 steps per launch (13% faster on C4 than 16 x 64)
            variant = 5 if self.dtype == torch.float32 else 1
        self.variant = int(variant)
        self.S = int(stage_steps)
        self.adjoint = bool(adjoint)
        self.graph = None
        self.adj_graph = None
        self._adj = None
        self._plans = {}
        self._seq = any(_has_media(p[0]) for p in problems)
        if self._seq:   # media: the single-problem path, variant by variant
            for coeffs, _, _ in problems:
                if coeffs.y_mirror_lo or coeffs.x_mirror_lo or coeffs.y_rot2_lo or coeffs.rot4_lo:
                    raise NotImplementedError("batched run: symmetry cuts are not supported")
            self.var = [{"coeffs": c, "source": s_, "monitors": list(m),
                         "eps": c.eps_inf.detach().to(self.dtype).clone().contiguous()} for c, s_, m in problems]
            return
        self.var = []
        for v, (coeffs, source, monitors) in enumerate(problems):
            self.var.append(self._alloc(coeffs, source, list(monitors)))
        # one descriptor workspace per plan (forward, and recompute with Ez tapes)
        self._ws_bytes = max(1, self.L.mt_tm2d_batch_workspace(self.B, self.code))
        self.ws = {}

    # ------------------------------------------------------------ allocation
    def _alloc(self, coeffs, source, monitors):
        if coeffs.y_mirror_lo or coeffs.x_mirror_lo or coeffs.y_rot2_lo or coeffs.rot4_lo:
            raise NotImplementedError("batched run: symmetry cuts are not supported")
        nx, ny = coeffs.eps_inf.shape
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=self.device)
        ny_e = ny if coeffs.y_pmc else ny - 2
        shapes = {"Ez": (nx, ny), "Dz": (nx, ny), "Hx": (nx, ny - 1), "Hy": (nx - 1, ny),
                  "psi_Ezy": (nx, ny - 1), "psi_Ezx": (nx - 1, ny), "psi_Hyx": (nx - 2, ny_e),
                  "psi_Hxy": (nx - 2, ny_e)}
        V = {"coeffs": coeffs, "nx": nx, "ny": ny, "monitors": monitors}
        V["eps"] = coeffs.eps_inf.detach().to(self.dtype).clone().contiguous()
        V["state"] = {f: z(*shapes[f]) for f in FIELDS}
        V["alt"] = {f: z(*shapes[f]) for f in FIELDS}
        nck = (self.n_steps + self.K - 1) // self.K
        V["ckpt"] = [{f: z(*shapes[f]) for f in FIELDS} for _ in range(nck)] if self.adjoint else []
        V["tape"] = z(self.K, nx, ny) if self.adjoint else None
        plan = engine2d.source_plan(source, (nx, ny), self.device, self.dtype,
                                    hshapes=engine2d.h_shapes("tm", nx, ny))
        V["plan"] = plan
        ev, hv = engine2d.coef_arrays(plan, coeffs.dt, 0, self.n_steps)
        # per-step dt * amplitude tables for the whole run, cast like the
        # single-problem path casts each step's coefficient
        tab = lambda arr: torch.tensor(list(arr)[:self.n_steps], dtype=torch.float64).to(self.dtype).to(self.device)
        V["coef"] = tab(ev) if ev is not None else None
        V["hcoef"] = tab(hv) if hv is not None else None
        arr, outs, keep = engine2d.monitor_structs(V["state"], monitors, self.K, self.dtype, self.device)
        V["mon_arr"], V["mon_keep"] = arr, keep
        V["chunk_acc"] = [o for (_, _, o) in outs]
        V["acc"] = [torch.zeros_like(o) for o in V["chunk_acc"]]
        V["outs_meta"] = outs
        if monitors:
            nbytes = self.L.mt_stage_bytes(arr, len(monitors), self.S, self.code)
            V["stage"] = torch.empty(max(1, nbytes), dtype=torch.uint8, device=self.device)
            V["stage_bytes"] = nbytes
        else:
            V["stage"], V["stage_bytes"] = None, 0
        return V

    def _prob(self, V, tape):
        c = V["coeffs"]
        d = E.TM2D()
        d.nx, d.ny, d.dtype, d.y_pmc = V["nx"], V["ny"], self.code, int(bool(c.y_pmc))
        d.dt, d.dx = float(c.dt), float(c.dx)
        d.eps = V["eps"].data_ptr()
        d.hy, d.hx, d.ex, d.ey = (engine2d._axis(c, n) for n in ("hy", "hx", "ex", "ey"))
        for f in FIELDS:
            setattr(d, f, V["state"][f].data_ptr())
        for q, f in enumerate(_ORDER):
            d.alt[q] = V["alt"][f].data_ptr()
        d.tape = V["tape"].data_ptr() if tape else 0
        d.variant = self.variant
        return d

    def _plan(self, tape):
        if tape in self._plans:
            return self._plans[tape]
        items = (BatchItem * self.B)()
        for v, V in enumerate(self.var):
            it = items[v]
            it.prob = self._prob(V, tape)
            s, _ = engine2d.source_struct(V["plan"])
            it.src = s
            it.coef_tab = V["coef"].data_ptr() if V["coef"] is not None else None
            if V["hcoef"] is not None:
                for slot, hb in V["plan"]["h"]:
                    engine2d._fill_source(it.prob.hsrc[slot], hb)
                it.hcoef_tab = V["hcoef"].data_ptr()
            it.mons = ctypes.cast(V["mon_arr"], ctypes.POINTER(E.Monitor))
            it.n_mons = len(V["monitors"])
            it.stage = V["stage"].data_ptr() if V["stage"] is not None else None
            it.stage_bytes = V["stage_bytes"]
        ws = self.ws.setdefault(tape, torch.empty(self._ws_bytes, dtype=torch.uint8, device=self.device))
        h = self.L.mt_tm2d_batch_create(items, self.B, self.variant, self.S, ws.data_ptr(), ws.numel(),
                                        E.stream_handle())
        if not h:
            E.check(-1, "mt_tm2d_batch_create")
        self._plans[tape] = (h, items)
        return self._plans[tape]

    def __del__(self):
        try:
            for h, _ in self._plans.values():
                self.L.mt_batch_destroy(h)
        except Exception:
            pass

    # ------------------------------------------------ media, variant by variant
    def _seq_run(self, V, eps):
        """One media variant through run_tm (differentiable in eps when eps
        requires grad and grad mode is on); returns its accumulators."""
        from .core.grid import Grid
        from .core.run import run_tm
        from .core.stepper_tm import init_tm_state
        c = V["coeffs"]._replace(eps_inf=eps)
        nx, ny = eps.shape
        g = Grid(nx=nx, ny=ny, dx=float(V["coeffs"].dx), dtype=self.dtype, device=self.device)
        mc = V["coeffs"].mc1x
        st = init_tm_state(g, n_poles=int(V["coeffs"].c1.shape[0]), y_pmc=bool(V["coeffs"].y_pmc),
                           n_mpoles=int(mc.shape[0]) if mc is not None else 0)
        _, accs = run_tm(st, c, self.n_steps, source=V["source"], monitors=V["monitors"],
                         checkpoint_every=self.K if self.adjoint else 0)
        return list(accs)

    # -------------------------------------------------------------- forward
    def set_eps(self, eps_list):
        with torch.no_grad():
            for V, e in zip(self.var, eps_list):
                V["eps"].copy_(e.detach().to(self.dtype))

    def _forward_body(self):
        h, _ = self._plan(False)
        for V in self.var:
            for t in V["state"].values():
                t.zero_()
            for a in V["acc"]:
                a.zero_()
        step0, ci = 0, 0
        while step0 < self.n_steps:
            k = min(self.K, self.n_steps - step0)
            for V in self.var:
                if self.adjoint:
                    for f in FIELDS:
                        V["ckpt"][ci][f].copy_(V["state"][f])
                for a in V["chunk_acc"]:
                    a.zero_()
            rc = self.L.mt_tm2d_batch_run(h, step0, k, E.stream_handle())
            E.check(rc, "mt_tm2d_batch_run")
            for V in self.var:
                for a, ca in zip(V["acc"], V["chunk_acc"]):
                    a.add_(ca)
            step0 += k
            ci += 1

    def capture(self):
        """Record the forward run in a HIP graph (replayed by forward())."""
        if self._seq:
            raise NotImplementedError("graph capture covers the batched launch, not variants with media")
        self._plan(False)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._forward_body()          # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._forward_body()
        self.graph = g

    def forward(self, eps_list=None, use_graph=True):
        """Run every variant; returns [[acc per monitor] per variant] shaped as
        the monitors' run_tm outputs (views of the fixed buffers).  With a
        captured graph the run replays it unless use_graph is False."""
        if eps_list is not None:
            self.set_eps(eps_list)
        if self._seq:
            with torch.no_grad():
                return [self._seq_run(V, V["eps"]) for V in self.var]
        if self.graph is not None and use_graph:
            self.graph.replay()
        else:
            self._forward_body()
        res = []
        for V in self.var:
            res.append([a.reshape(m.out_shape(V["state"])) for a, m in zip(V["acc"], V["monitors"])])
        return res

    # -------------------------------------------------------------- adjoint
    def _adj_setup(self):
        """Persistent adjoint buffers (cotangent fields, transposed-difference
        scratch, eps gradients, monitor cotangents) and their descriptors,
        uploaded once (mt_tm2d_batch_adjoint_prepare): every later sweep
        only launches, so it can be replayed from a graph."""
        if self._adj is not None:
            return self._adj
        if not self.adjoint:
            raise RuntimeError("BatchedTM was built with adjoint=False (no checkpoints or tapes)")
        from .core.adjoint2d import _adj_monitors
        h, _ = self._plan(True)
        B = self.B
        adjs, tapes = (E.TM2DAdj * B)(), (E.c_void_p * B)()
        per, mons, off, keep = [], [], [0], []
        for v, V in enumerate(self.var):
            nx, ny = V["nx"], V["ny"]
            A = {f: torch.zeros_like(V["state"][f]) for f in FIELDS}
            sc = {"gHy": torch.empty(nx, ny, dtype=self.dtype, device=self.device),
                  "gHx": torch.empty(nx, ny, dtype=self.dtype, device=self.device),
                  "gEx": torch.empty(nx - 1, ny, dtype=self.dtype, device=self.device),
                  "gEy": torch.empty(nx, ny - 1, dtype=self.dtype, device=self.device)}
            g = torch.zeros(nx, ny, dtype=self.dtype, device=self.device)
            gacc = [torch.zeros_like(a) for a in V["acc"]]
            a = adjs[v]
            for f in FIELDS:
                setattr(a, "a" + f, A[f].data_ptr())
            for f, t in sc.items():
                setattr(a, f, t.data_ptr())
            a.eps_grad = g.data_ptr()
            tapes[v] = V["tape"].data_ptr()
            arr, n_used, kp = _adj_monitors(A, V["monitors"], gacc, self.dtype, self.device)
            keep.append(kp)
            for i in range(n_used):
                mons.append(arr[i])
            off.append(off[-1] + n_used)
            per.append({"A": A, "sc": sc, "g": g, "gacc": gacc})
        mon_arr = (E.Monitor * max(1, len(mons)))(*mons)
        off_arr = (E.c_int * (B + 1))(*off)
        rc = self.L.mt_tm2d_batch_adjoint_prepare(h, adjs, tapes, mon_arr, off_arr, E.stream_handle())
        E.check(rc, "mt_tm2d_batch_adjoint_prepare")
        # second cotangent set: the fused reverse step reads one set and
        # writes the other (batched variants carry no conductivity or poles)
        fused = E.c_int(0)
        adjs2, mons2, keep2 = (E.TM2DAdj * B)(), [], []
        if self.fused_adjoint:
            for v, (V, P_) in enumerate(zip(self.var, per)):
                A2 = {f: torch.zeros_like(t) for f, t in P_["A"].items()}
                a2 = adjs2[v]
                for f in FIELDS:
                    setattr(a2, "a" + f, A2[f].data_ptr())
                for f, t in P_["sc"].items():
                    setattr(a2, f, t.data_ptr())
                a2.eps_grad = P_["g"].data_ptr()
                arr, n_used, kp = _adj_monitors(A2, V["monitors"], P_["gacc"], self.dtype, self.device)
                keep2.append(kp)
                mons2.extend(arr[i] for i in range(n_used))
                P_["A2"] = A2
            mon_arr2 = (E.Monitor * max(1, len(mons2)))(*mons2)
            rc = self.L.mt_tm2d_batch_adjoint_pingpong(h, adjs2, mon_arr2, off_arr, ctypes.byref(fused),
                                                       E.stream_handle())
            E.check(rc, "mt_tm2d_batch_adjoint_pingpong")
            keep2.append((adjs2, mon_arr2))
        self._adj = {"h": h, "per": per, "fused": bool(fused.value),
                     "keep": (adjs, tapes, mon_arr, off_arr, keep, keep2)}
        return self._adj

    def _gradient_body(self):
        adj = self._adj
        h = adj["h"]
        for P in adj["per"]:
            for t in P["A"].values():
                t.zero_()
            P["g"].zero_()
        nck = len(self.var[0]["ckpt"])
        for ci in reversed(range(nck)):
            step0 = ci * self.K
            k = min(self.K, self.n_steps - step0)
            for V in self.var:
                for f in FIELDS:
                    V["state"][f].copy_(V["ckpt"][ci][f])
            rc = self.L.mt_tm2d_batch_run(h, step0, k, E.stream_handle())
            E.check(rc, "mt_tm2d_batch_run (recompute)")
            rc = self.L.mt_tm2d_batch_adjoint_sweep(h, step0, k, E.stream_handle())
            E.check(rc, "mt_tm2d_batch_adjoint_sweep")

    def gradient(self, gaccs):
        """dL/deps per variant given dL/dacc (list per variant of per-monitor
        cotangents, None where unused).  Uses the checkpoints of the last
        forward; replays the captured adjoint graph when there is one.
        Media variants run their forward again through the single-problem
        adjoint (run_tm's checkpointed reverse sweep)."""
        if self._seq:
            if not self.adjoint:
                raise RuntimeError("BatchedTM was built with adjoint=False (no checkpoints or tapes)")
            out = []
            for V, gv in zip(self.var, gaccs):
                e = V["eps"].clone().requires_grad_(True)
                with torch.enable_grad():
                    accs = self._seq_run(V, e)
                    pairs = [(a, g_) for a, g_ in zip(accs, gv) if g_ is not None]
                    if not pairs:
                        out.append(torch.zeros_like(e))
                        continue
                    (ge,) = torch.autograd.grad([a for a, _ in pairs],
                                                [e], [g_.reshape(a.shape).to(a.dtype) for a, g_ in pairs])
                out.append(ge.detach())
            return out
        adj = self._adj_setup()
        with torch.no_grad():
            for P, gv in zip(adj["per"], gaccs):
                for buf, g in zip(P["gacc"], gv):
                    if g is None:
                        buf.zero_()
                    else:
                        buf.copy_(g.detach().reshape(buf.shape))
        if self.adj_graph is not None:
            self.adj_graph.replay()
        else:
            self._gradient_body()
        return [P["g"].clone() for P in adj["per"]]

    def capture_adjoint(self):
        """Record the reverse sweep (checkpoint restores, chunk recomputes on
        the Ez tapes, batched adjoint steps) in a HIP graph that gradient()
        replays; the cotangents are copied into fixed buffers before each
        replay."""
        if self._seq:
            raise NotImplementedError("graph capture covers the batched launch, not variants with media")
        self._adj_setup()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._gradient_body()         # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._gradient_body()
        self.adj_graph = g

    def apply(self, eps_list):
        """Differentiable forward: returns [[acc per monitor] per variant];
        backward gives dL/deps for each variant's eps tensor."""
        if self._seq:
            self.set_eps(eps_list)
            return [self._seq_run(V, e) for V, e in zip(self.var, eps_list)]
        flat = _BatchFn.apply(self, *eps_list)
        res, i = [], 0
        for V in self.var:
            n = len(V["monitors"])
            res.append(list(flat[i:i + n]))
            i += n
        return res


class _BatchFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, runner, *eps_list):
        out = runner.forward(list(eps_list))
        ctx.runner = runner
        ctx.counts = [len(o) for o in out]
        return tuple(a.clone() for o in out for a in o)

    @staticmethod
    def backward(ctx, *g):
        runner = ctx.runner
        gaccs, i = [], 0
        for n in ctx.counts:
            gaccs.append([x.contiguous() if x is not None else None for x in g[i:i + n]])
            i += n
        grads = runner.gradient(gaccs)
        return (None, *grads)
