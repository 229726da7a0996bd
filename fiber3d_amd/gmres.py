"""Right-preconditioned restarted GMRES in FP64 on the GPU.

Semantics of bpm_utilities/gmres.py:94-108 over SciPy's gmres as called at
python/tstep/tstep.py:1209-1216: solve (A P^-1) y = b, return x = P^-1 y,
stop when the residual norm <= tol ||b|| (atol = 0), restart length and
maximum restart cycles as given.  An initial guess x0 is taken in x-space
(r0 = b - A x0), so a predicted solution can seed the solve; a guess that is
not finite, or whose residual is larger than ||b||, is dropped and the solve
starts from zero (the plain method stays the fallback).

Arnoldi uses classical Gram-Schmidt with one re-orthogonalisation (CGS2) as
two matrix-vector products with the basis, all on the device in FP64.  The
(restart+1) x restart Hessenberg least-squares problem is updated with Givens
rotations on the host (j+2 scalars per iteration).

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


def gmres(matvec, b, precond=None, tol=1e-8, restart=150, maxiter=200, x0=None, inner=None,
          callback=None, fused_inner=None, fused_matvec=None, update_residual=False,
          inner_rel_err=1e-6, target=None, measure_inner=False):
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
                               inner_cycles, "fp64", beta / bnorm, measured)
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
                                   inner_cycles + 1, "update", bound / bnorm, measured)
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
                       inner_cycles, "fp64", rn, measured)
