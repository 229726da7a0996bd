"""Right-preconditioned restarted GMRES in FP64 on the GPU.

Semantics of bpm_utilities/gmres.py:94-108 over SciPy's gmres as called at
python/tstep/tstep.py:1209-1216: solve (A P^-1) y = b, return x = P^-1 y,
stop when the residual norm <= tol ||b|| (atol = 0), restart length and
maximum restart cycles as given.  An initial guess x0 is taken in x-space
(r0 = b - A x0), so a predicted solution can seed the solve; a guess that is
not finite, or whose residual is larger than ||b||, is dropped and the solve
starts from zero (the plain method stays the fallback).

Arnoldi uses classical Gram-Schmidt with one re-orthogonalisation (CGS2) in
FP64.  Two engines run the Arnoldi steps and the (restart+1) x restart
Hessenberg least-squares problem:
  * "device" (for GPU tensors): libf3d's GMRES kernels (f3d_gmres_*) keep the
    basis, the Hessenberg matrix, its Givens rotations and the residual
    estimate in device memory; an Arnoldi step is a fixed sequence of kernel
    launches with the step index in device memory, so a GmresContext can
    record the whole step (operator included) as one HIP graph.  The host
    reads the residual estimates through an asynchronous copy: `expect`
    steps run back to back before the first read (the caller's estimate of
    the iteration count, e.g. the previous solve's), then one read per step;
  * "host": torch products with the basis and Givens rotations in NumPy, one
    read per step; it runs on any device (the algorithm's CPU tests).

An optional `inner` operator (e.g. the FP32-compensated matvec) is used for
the Arnoldi products while the true residual that decides convergence and
restarts is always recomputed with the FP64 `matvec`.  The inner operator's
error is relative to the size of the vector it acts on, so it limits how far
one cycle can reduce the true residual: when a cycle converges by its own
estimate but the FP64 residual is still above tol, the inner operator has
reached its accuracy floor and the remaining cycles use `matvec`.

`fused_inner` / `fused_matvec`, when given, compute v -> A P^-1 v in one call
(e.g. a recorded HIP graph of the preconditioner and the operator) and replace
the separate precond + operator calls of the Arnoldi step.

`update_residual` (with an inner operator) replaces the FP64 product that
ends an inner cycle by an update of the cycle's FP64 starting residual r:
the correction's image under the inner operator is V_{k+1} H y by the Arnoldi
relation, so r_new = r - V_{k+1} (H y) costs one small GEMV with the basis.
It differs from the true b - A x_new only by the inner operator's error on
the correction, at most `inner_rel_err` * ||H y||.  The cycle is accepted as
converged when ||r_new|| + inner_rel_err * ||H y|| <= tol ||b||; otherwise
the FP64 residual is computed as usual.  From a zero guess ||H y|| ~ ||b||,
so the bound never passes and the FP64 product is kept; from a close guess
||H y|| ~ ||b - A x0||, and one FP64 product per solve (the guess's residual)
remains.

Whenever an inner cycle ends with the FP64 residual instead, the same update
is set beside it.  Their difference, sum_j y_j A_inner P^-1 v_j -
A P^-1 (sum_j y_j v_j), is exactly what the update misses: the inner
operator's error on the correction plus the rounding of a low-precision P^-1
applied to the combination rather than to each basis vector.
`measure_inner=True` returns the largest such difference relative to ||H y||
as `inner_err`, the quantity `inner_rel_err` must bound.
"""
import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class GmresResult:
    x: torch.Tensor
    info: int
    iterations: int
    restarts: int
    residual: float          # true ||b - A x|| / ||b|| in FP64
    guess_used: bool
    history: list
    inner_cycles: int = 0    # restart cycles run with the inner operator
    residual_kind: str = "fp64"  # "fp64": b - A x recomputed; "update": FP64-anchored update
    residual_bound: float = 0.0  # bound on the true relative residual
    inner_err: float = None      # measured ||(A_inner - A) z|| / ||A_inner z|| (max over cycles)
    resid: torch.Tensor = None   # the residual vector the closing decision used (b - A x,
                                 # or the FP64-anchored update)


class GmresContext:
    """Device buffers of the "device" engine for systems of size n and restart
    m: the basis V ((m+1) x n), the operator's input and output vectors and
    the solver state (Hessenberg, rotations, estimates, counters), allocated
    once and reused by every solve; and the recorded HIP graphs of whole
    Arnoldi steps (the operator's input is gathered from V on the device, so
    one graph serves every step of every cycle).  A graph addresses the
    buffers of everything its operator reads: drop it (reset) when those
    change."""

    def __init__(self, n, m, device):
        from . import _lib as L
        self.L = L
        self.n, self.m = int(n), int(m)
        if not 1 <= self.m <= 159:
            raise ValueError("restart must be in [1, 159] for the device engine")
        lib = L.load()
        nbytes = lib.f3d_gmres_workspace_size(self.n, self.m)
        self.ws = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        self.V = torch.zeros((self.m + 1, self.n), dtype=torch.float64, device=device)
        self.xin = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.z = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.w = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.upd = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.r_upd = torch.zeros(self.n, dtype=torch.float64, device=device)
        # state views: est[k] after step k, the break flag, the step index
        off = self._layout()
        self.est = self.ws[off["est"]:off["est"] + 8 * self.m].view(torch.float64)
        self.ctl = self.ws[off["ctl"]:off["ctl"] + 32].view(torch.int32)
        self.hy = self.ws[off["y"] + 8 * (self.m + 1):off["y"] + 16 * (self.m + 1)].view(torch.float64)
        self.est_h = torch.zeros(self.m + 1, dtype=torch.float64, pin_memory=True)
        self.brk_h = torch.zeros(self.m + 1, dtype=torch.int32, pin_memory=True)
        self.events = [torch.cuda.Event() for _ in range(self.m + 1)]
        self.graphs = {}
        self.captures = 0

    def _layout(self):
        # byte offsets of gmres_layout (fiber3d_amd/csrc/gmres.hip)
        n, m = self.n, self.m
        nch = min((n + 4095) // 4096, 1024)
        off, o = {}, 0
        def take(name, count):
            nonlocal o
            off[name] = o
            o += (count * 8 + 255) // 256 * 256
        take("H", (m + 1) * m)
        take("Hr", (m + 1) * m)
        take("cs", m)
        take("sn", m)
        take("g", m + 1)
        take("est", m)
        take("h", m + 1)
        take("part", max((m + 1) * nch, 2048))
        take("y", 2 * (m + 1))
        take("ctl", 8)
        assert o == self.ws.numel(), (o, self.ws.numel())
        return off

    def reset(self):
        self.graphs = {}

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def begin(self, r):
        L = self.L
        L.check(L.load().f3d_gmres_begin(self.n, self.m, L.ptr(r), L.ptr(self.V), L.ptr(self.ws),
                                         self._stream()), "f3d_gmres_begin")

    def _step_eager(self, op):
        L, lib = self.L, self.L.load()
        L.check(lib.f3d_gmres_current(self.n, self.m, L.ptr(self.V), L.ptr(self.xin),
                                      L.ptr(self.ws), self._stream()), "f3d_gmres_current")
        op(self.xin, self.w)
        L.check(lib.f3d_gmres_arnoldi(self.n, self.m, L.ptr(self.V), L.ptr(self.w),
                                      L.ptr(self.ws), self._stream()), "f3d_gmres_arnoldi")

    def step(self, op, key=None):
        """One Arnoldi step with op(x, out) = A P^-1 x; replayed from a
        recorded graph when key is given (recorded at its first use)."""
        if key is None:
            self._step_eager(op)
            return
        g = self.graphs.get(key)
        if g is None:
            # the first step runs eagerly (warm-up: every buffer the operator
            # touches is allocated by then); the graph is recorded on a
            # side stream from the next one on
            self._step_eager(op)
            self.graphs[key] = "warm"
            return
        if g == "warm":
            torch.cuda.current_stream().synchronize()
            ctl = self.ctl.clone()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self._step_eager(op)
            # recording ran nothing: replay it for this step
            self.ctl.copy_(ctl)
            self.graphs[key] = g = graph
            self.captures += 1
        g.replay()

    def post(self, j):
        """Queue the read-back of step j's estimate and break flag."""
        self.est_h[j:j + 1].copy_(self.est[j:j + 1], non_blocking=True)
        self.brk_h[j:j + 1].copy_(self.ctl[1:2], non_blocking=True)
        self.events[j].record()

    def read(self, j):
        self.events[j].synchronize()
        return float(self.est_h[j]), int(self.brk_h[j])

    def solve(self, k, r=None):
        """upd = V_k y (y: the least-squares solution of the first k steps);
        with r, also r_upd = r - V_{k+1} (Hr y)."""
        L = self.L
        L.check(L.load().f3d_gmres_solve(self.n, self.m, int(k), L.ptr(self.V), L.ptr(self.upd),
                                         L.ptr(r), L.ptr(self.r_upd if r is not None else None),
                                         L.ptr(self.ws), self._stream()), "f3d_gmres_solve")
        return self.upd


def gmres(matvec, b, precond=None, tol=1e-8, restart=150, maxiter=200, x0=None, inner=None,
          callback=None, fused_inner=None, fused_matvec=None, update_residual=False,
          inner_rel_err=1e-6, target=None, measure_inner=False, engine=None, ctx=None,
          expect=0, graph_keys=None):
    """engine: "device" (default for GPU tensors) or "host" (see the module
    docstring); ctx: a GmresContext to reuse (device engine); expect: steps
    to run before the first read of the residual estimate; graph_keys: dict
    {operator: key} (operator is matvec or inner) naming the recorded graphs
    of ctx to replay for Arnoldi steps with that operator."""
    if engine is None:
        engine = "device" if b.is_cuda else "host"
    if engine == "device":
        return _gmres_device(matvec, b, precond, tol, restart, maxiter, x0, inner, callback,
                             fused_inner, fused_matvec, update_residual, inner_rel_err, target,
                             measure_inner, ctx, expect, graph_keys)
    if engine != "host":
        raise ValueError("engine is 'device' or 'host'")
    return _gmres_host(matvec, b, precond, tol, restart, maxiter, x0, inner, callback,
                       fused_inner, fused_matvec, update_residual, inner_rel_err, target,
                       measure_inner)


def _gmres_host(matvec, b, precond, tol, restart, maxiter, x0, inner, callback, fused_inner,
                fused_matvec, update_residual, inner_rel_err, target, measure_inner):
    # iterations continue to `target` (<= tol) once started; a guess (or any
    # restart) whose FP64 residual is already <= tol is accepted as it is
    t_it = tol if target is None else min(float(target), tol)
    Pinv = precond if precond is not None else (lambda v, out=None: v.clone())
    op = inner if inner is not None else matvec
    inner_cycles = 0
    measured = None
    n = b.numel()
    dev = b.device
    bnorm = float(torch.linalg.norm(b))
    hist = []
    if bnorm == 0.0:
        return GmresResult(torch.zeros_like(b), 0, 0, 0, 0.0, False, hist, 0)
    guess_used = False
    x = torch.zeros_like(b)
    r = b.clone()
    if x0 is not None:
        if bool(torch.isfinite(x0).all()):
            r0 = b - matvec(x0)
            if float(torch.linalg.norm(r0)) < bnorm:
                x = x0.clone()
                r = r0
                guess_used = True
    m = min(restart, n)
    V = torch.empty((m + 1, n), dtype=torch.float64, device=dev)
    total = 0
    restarts = 0
    w = torch.empty_like(b)
    z = torch.empty_like(b)
    for _cycle in range(maxiter):
        beta = float(torch.linalg.norm(r))
        hist.append(beta / bnorm)
        if beta <= tol * bnorm:
            return GmresResult(x, 0, total, restarts, beta / bnorm, guess_used, hist,
                               inner_cycles, "fp64", beta / bnorm, measured, r)
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        Hraw = np.zeros((m + 1, m))  # unrotated Hessenberg: A P^-1 V_k = V_{k+1} Hraw
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        j_done = 0
        fused = fused_matvec if op is matvec else fused_inner
        for j in range(m):
            if fused is not None:
                fused(V[j], out=w)
            else:
                Pinv(V[j], out=z)
                op(z, out=w)
            # CGS2 against V[0..j]
            Vj = V[:j + 1]
            h1 = torch.mv(Vj, w)
            w -= torch.mv(Vj.t(), h1)
            h2 = torch.mv(Vj, w)
            w -= torch.mv(Vj.t(), h2)
            hn = torch.linalg.norm(w)
            hcol = torch.cat((h1 + h2, hn.reshape(1))).cpu().numpy()
            H[:j + 2, j] = hcol
            Hraw[:j + 2, j] = hcol
            if hcol[j + 1] > 0:
                V[j + 1] = w / hcol[j + 1]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (H[j, j] / den, H[j + 1, j] / den) if den > 0 else (1.0, 0.0)
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total += 1
            j_done = j + 1
            est = abs(g[j + 1]) / bnorm
            hist.append(est)
            if callback is not None:
                callback(est)
            if est <= t_it or hcol[j + 1] == 0.0:
                break
        yk = np.linalg.solve(np.triu(H[:j_done, :j_done]), g[:j_done]) if j_done else np.zeros(0)
        upd = torch.mv(V[:j_done].t(), torch.from_numpy(yk).to(dev))
        x += Pinv(upd)
        restarts += 1
        r_upd = None
        if (update_residual or measure_inner) and op is not matvec and j_done:
            hy = Hraw[:j_done + 1, :j_done] @ yk
            hyn = float(np.linalg.norm(hy))
            k1 = j_done + 1 if hy[j_done] != 0.0 else j_done  # V[j_done] unset at breakdown
            r_upd = r - torch.mv(V[:k1].t(), torch.from_numpy(hy[:k1]).to(dev))
            rn_upd = float(torch.linalg.norm(r_upd))
            bound = rn_upd + inner_rel_err * hyn
            if update_residual and est <= t_it and bound <= tol * bnorm:
                hist.append(rn_upd / bnorm)
                return GmresResult(x, 0, total, restarts, rn_upd / bnorm, guess_used, hist,
                                   inner_cycles + 1, "update", bound / bnorm, measured,
                                   r_upd.clone())
        r = b - matvec(x)
        if r_upd is not None and hyn > 0.0:
            # r_old - r_upd is the correction's image under the inner
            # operator as the Arnoldi relation gives it, r_old - r its FP64
            # image: the difference is the update's whole error, measured at
            # no extra cost
            e = float(torch.linalg.norm(r - r_upd)) / hyn
            measured = e if measured is None else max(measured, e)
        if op is not matvec:
            inner_cycles += 1
            if est <= t_it and float(torch.linalg.norm(r)) > tol * bnorm:
                op = matvec
    rn = float(torch.linalg.norm(r)) / bnorm
    return GmresResult(x, 0 if rn <= tol else total, total, restarts, rn, guess_used, hist,
                       inner_cycles, "fp64", rn, measured, r)


def _gmres_device(matvec, b, precond, tol, restart, maxiter, x0, inner, callback, fused_inner,
                  fused_matvec, update_residual, inner_rel_err, target, measure_inner, ctx,
                  expect, graph_keys):
    """The "device" engine: the same algorithm as _gmres_host, with the
    Arnoldi steps and the Hessenberg least-squares problem on the device
    (GmresContext); the host reads the residual estimate of each step
    through an asynchronous copy and the norms once per cycle."""
    t_it = tol if target is None else min(float(target), tol)

    def Pinv(v, out=None):
        if precond is None:
            return v.clone() if out is None else out.copy_(v)
        return precond(v, out=out)
    op = inner if inner is not None else matvec
    inner_cycles = 0
    measured = None
    n = b.numel()
    hist = []
    guess_used = False
    beta0 = None
    if x0 is not None:
        # |b| and the guess's residual |b - A x0| in one read: a non-finite
        # guess gives a non-finite residual and is dropped like a poor one
        r0 = b - matvec(x0)
        bnorm, r0n = torch.stack([torch.linalg.norm(b), torch.linalg.norm(r0)]).tolist()
    else:
        bnorm, r0n = float(torch.linalg.norm(b)), float("nan")
    if bnorm == 0.0:
        return GmresResult(torch.zeros_like(b), 0, 0, 0, 0.0, False, hist, 0)
    if math.isfinite(r0n) and r0n < bnorm:
        x = x0.clone()
        r = r0
        guess_used = True
        beta0 = r0n
    else:
        x = torch.zeros_like(b)
        r = b.clone()
        beta0 = bnorm
    m = min(restart, n)
    if ctx is None or ctx.n != n or ctx.m != m:
        ctx = GmresContext(n, m, b.device)
    keys = graph_keys or {}
    total = restarts = 0
    ahead = max(0, int(expect))
    for _cycle in range(maxiter):
        beta = beta0 if beta0 is not None else float(torch.linalg.norm(r))
        beta0 = None
        hist.append(beta / bnorm)
        if beta <= tol * bnorm:
            return GmresResult(x, 0, total, restarts, beta / bnorm, guess_used, hist,
                               inner_cycles, "fp64", beta / bnorm, measured, r)
        ctx.begin(r)
        fused = fused_matvec if op is matvec else fused_inner
        if fused is not None:
            opf = fused
        else:
            def opf(v, out, op=op):
                op(Pinv(v, out=ctx.z), out=out)
        key = keys.get("matvec" if op is matvec else "inner")
        j = j_read = 0
        k = None
        est = 1.0
        while k is None:
            ctx.step(opf, key)
            ctx.post(j)
            j += 1
            if j < min(ahead, m):
                continue
            while j_read < j and k is None:
                e, brk = ctx.read(j_read)
                j_read += 1
                est = e / bnorm
                hist.append(est)
                if callback is not None:
                    callback(est)
                if est <= t_it or brk:
                    k = j_read
            if k is None and j == m:
                k = m
        ahead = 0  # later cycles read every step
        total += k
        need_rupd = (update_residual or measure_inner) and op is not matvec
        upd = ctx.solve(k, r if need_rupd else None)
        x += Pinv(upd)
        restarts += 1
        r_upd = None
        if need_rupd:
            hyn = float(torch.linalg.norm(ctx.hy[:k + 1]))
            r_upd = ctx.r_upd
            rn_upd = float(torch.linalg.norm(r_upd))
            bound = rn_upd + inner_rel_err * hyn
            if update_residual and est <= t_it and bound <= tol * bnorm:
                hist.append(rn_upd / bnorm)
                return GmresResult(x, 0, total, restarts, rn_upd / bnorm, guess_used, hist,
                                   inner_cycles + 1, "update", bound / bnorm, measured,
                                   r_upd.clone())
        r = b - matvec(x)
        if r_upd is not None and hyn > 0.0:
            # the update's whole error, as in _gmres_host
            e = float(torch.linalg.norm(r - r_upd)) / hyn
            measured = e if measured is None else max(measured, e)
        if op is not matvec:
            inner_cycles += 1
            if est <= t_it and float(torch.linalg.norm(r)) > tol * bnorm:
                op = matvec
    rn = float(torch.linalg.norm(r)) / bnorm
    return GmresResult(x, 0 if rn <= tol else total, total, restarts, rn, guess_used, hist,
                       inner_cycles, "fp64", rn, measured, r)
