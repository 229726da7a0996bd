"""Fixed-dt implicit time stepping of free fibers in a shell on the GPU.

One step follows tstep.time_step_hydro (python/tstep/tstep.py:628-1216):
assemble (FiberSystem.prepare), solve with right-preconditioned GMRES to the
relative tolerance, then update fiber positions, tensions and the shell
density from the solution (tstep.py:1389-1397).  Free rigid bodies move
rigidly with their solved velocities.  Adaptive step rejection, dynamic
instability and output writing of the reference driver are not part of this
path.

Two modes:
  'plain'  : FP64 pair sums, FP64 fiber LU, zero initial guess (the reference
             method on the GPU);
  'fast'   : FP32-compensated pair sums inside the Krylov iteration with the
             FP64 operator deciding convergence, FP32 fiber LU with FP64
             refinement, and an optional initial-guess predictor.
"""
import time
from dataclasses import dataclass

import numpy as np
import torch

from .gmres import gmres
from .graphs import ArnoldiGraphs
from .system import FiberSystem


@dataclass
class StepStats:
    iterations: int
    restarts: int
    residual: float
    seconds: float
    guess_used: bool


class TimeStepper:
    def __init__(self, fibers, shell, mode="plain", eta=1.0, dt=1e-3, E=0.01, epsilon=1e-3,
                 penalty=500.0, nfd=4, end_force=1.0, discretisation="reference", tol=1e-8,
                 restart=150, maxiter=200, predictor=None, lengths=None, bodies=None,
                 body_force=None, graph=False):
        if mode not in ("plain", "fast"):
            raise ValueError("mode is 'plain' or 'fast'")
        self.mode = mode
        self.tol, self.restart, self.maxiter = tol, restart, maxiter
        fast = mode == "fast"
        self.sys = FiberSystem(fibers, shell, eta=eta, dt=dt, E=E, epsilon=epsilon,
                               penalty=penalty, nfd=nfd, end_force=end_force,
                               discretisation=discretisation, lengths=lengths,
                               bodies=bodies, body_force=body_force,
                               # the system's own operator (true residual,
                               # convergence) is always FP64; the fast mode
                               # adds the FP32-compensated inner operator
                               pair_precision="f64",
                               lu_precision="f32c" if fast else "f64",
                               # a preconditioner need not be exact: the FP32
                               # factors are applied without refinement (a fixed
                               # linear map), the FP64 residual judges convergence
                               lu_refine=0)
        self.predictor = predictor
        S = self.sys
        # recorded A P^-1 per operator precision, re-recorded after each assembly
        self.graphs = ArnoldiGraphs(S) if graph else None
        dev = S.dev
        # solution index of every node coordinate: X(N), Y(N), Z(N), T(N) per fiber
        idx = np.empty((S.npts, 3), dtype=np.int64)
        tidx = np.empty(S.npts, dtype=np.int64)
        for k in range(S.nfib):
            N = S.sizes[k]
            base = S.fib0 + 4 * S.offs[k]
            for c in range(3):
                idx[S.offs[k]:S.offs[k + 1], c] = base + c * N + np.arange(N)
            tidx[S.offs[k]:S.offs[k + 1]] = base + 3 * N + np.arange(N)
        self.idx_x = torch.from_numpy(idx).to(dev)
        self.idx_t = torch.from_numpy(tidx).to(dev)
        self.tension = torch.zeros(S.npts, dtype=torch.float64, device=dev)
        self.shell_density = torch.zeros(3 * S.nper, dtype=torch.float64, device=dev)
        self.prev_solution = None
        self.history = []

    def inner_operator(self):
        S = self.sys
        if self.mode == "fast":
            return lambda v, out=None: S.matvec(v, out=out, precision="f32c")
        return None

    def step(self):
        S = self.sys
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x0 = None
        if self.predictor is not None:
            x0 = self.predictor.predict(self)
        fused = {}
        if self.graphs is not None:
            G = self.graphs
            fused["fused_matvec"] = lambda v, out: G.get("f64")(v, out=out)
            if self.mode == "fast":
                fused["fused_inner"] = lambda v, out: G.get("f32c")(v, out=out)
        res = gmres(S.matvec, S.rhs, S.precond, tol=self.tol, restart=self.restart,
                    maxiter=self.maxiter, x0=x0, inner=self.inner_operator(), **fused)
        sol = res.x
        if self.predictor is not None:
            self.predictor.observe(self, sol)
        self.prev_solution = sol
        self.shell_density = sol[:3 * S.nper].clone()
        self.tension = sol[self.idx_t]
        S.X.copy_(sol[self.idx_x])
        if S.bodies is not None:
            # rigid motion with the solved velocities (tstep.py:1398-1406)
            S.bodies.move(S.bodies._vel(sol, 3 * S.nper), S.dt)
        # re-assembly refills the system's buffers in place: the recorded
        # Arnoldi graphs stay valid
        S.prepare()
        torch.cuda.synchronize()
        st = StepStats(res.iterations, res.restarts, res.residual, time.perf_counter() - t0,
                       res.guess_used)
        self.history.append(st)
        return st
