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

In fast mode the closing residual is, by default (certify="update"), the
guess's FP64 residual updated by the correction's image under the
FP32-compensated operator, accepted only under a bound that includes that
operator's error (fiber3d_amd.gmres); certify="fp64" recomputes b - A x in
FP64 instead.  With record=True every step keeps its geometry, right-hand
side and solution, and verify() re-assembles each recorded step and returns
the true FP64 residuals.
"""
import time
from dataclasses import dataclass

import numpy as np
import torch

from .fileio import write_fibers_frame, write_resume_fibers
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
    residual_kind: str = "fp64"
    residual_bound: float = 0.0


class TimeStepper:
    def __init__(self, fibers, shell, mode="plain", eta=1.0, dt=1e-3, E=0.01, epsilon=1e-3,
                 penalty=500.0, nfd=4, end_force=1.0, discretisation="reference", tol=1e-8,
                 restart=150, maxiter=200, predictor=None, lengths=None, bodies=None,
                 body_force=None, graph=False, certify="update", record=False, lu_every=None,
                 target=None):
        if mode not in ("plain", "fast"):
            raise ValueError("mode is 'plain' or 'fast'")
        if certify not in ("update", "fp64"):
            raise ValueError("certify is 'update' or 'fp64'")
        self.mode = mode
        self.certify = certify
        self.record = record
        self.records = []
        # fiber LU factors are recomputed every lu_every-th assembly and
        # reused as a lagged preconditioner in between (fast mode: every 8th)
        self.lu_every = max(1, int(lu_every if lu_every is not None
                                   else (8 if mode == "fast" else 1)))
        self.target = target
        self.nsteps = 0
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
                    maxiter=self.maxiter, x0=x0, inner=self.inner_operator(),
                    update_residual=self.mode == "fast" and self.certify == "update",
                    target=self.target, **fused)
        sol = res.x
        if self.record:
            # S.rhs and sol are fresh tensors of this step; X is updated in place
            self.records.append((S.X.clone(), S.bodies.state() if S.bodies is not None else None,
                                 S.rhs, sol))
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
        self.nsteps += 1
        S.prepare(refactor=self.nsteps % self.lu_every == 0)
        torch.cuda.synchronize()
        st = StepStats(res.iterations, res.restarts, res.residual, time.perf_counter() - t0,
                       res.guess_used, res.residual_kind, res.residual_bound)
        self.history.append(st)
        return st

    def fiber_arrays(self):
        """Host copies of every fiber's positions (N_k, 3) and tensions (N_k,)."""
        S = self.sys
        X = S.X.cpu().numpy()
        T = self.tension.cpu().numpy()
        o = S.offs
        return ([X[o[k]:o[k + 1]] for k in range(S.nfib)],
                [T[o[k]:o[k + 1]] for k in range(S.nfib)])

    def write_frame(self, f, txt=True):
        """Append the current fibers as one frame of the reference's
        <output>_<ID>_fibers.txt (txt=True) or <output>_<ID>.fibers output
        (tstep.py:2158-2177); f is opened in binary mode."""
        X, T = self.fiber_arrays()
        write_fibers_frame(f, X, T, self.sys.E, self.sys.lengths_h, txt=txt)

    def write_resume(self, name_file):
        """The reference's <output>_<ID>_resume.fibers for the current fibers."""
        X, _ = self.fiber_arrays()
        write_resume_fibers(name_file, X, self.sys.E, self.sys.lengths_h)

    def verify(self):
        """True relative residuals ||b - A x|| / ||b|| of the recorded steps in
        FP64, each with the system re-assembled at that step's geometry, and
        the largest difference between the re-assembled and the recorded
        right-hand sides.  The current state is restored afterwards."""
        S = self.sys
        now = (S.X.clone(), S.bodies.state() if S.bodies is not None else None)
        out, drhs = [], 0.0
        for X, body, b, x in self.records:
            S.X.copy_(X)
            if body is not None:
                S.bodies.set_state(body)
            S.prepare()
            drhs = max(drhs, float((S.rhs - b).abs().max()))
            r = b - S.matvec(x, precision="f64")
            out.append(float(torch.linalg.norm(r) / torch.linalg.norm(b)))
        S.X.copy_(now[0])
        if now[1] is not None:
            S.bodies.set_state(now[1])
        S.prepare()
        return out, drhs
