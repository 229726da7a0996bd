"""Fixed-dt implicit time stepping of free fibers in a shell on the GPU.

One step follows tstep.time_step_hydro (python/tstep/tstep.py:628-1216):
assemble (FiberSystem.prepare), solve with right-preconditioned GMRES to the
relative tolerance, then update fiber positions, tensions and the shell
density from the solution (tstep.py:1389-1397).  Free rigid bodies move
rigidly with their solved velocities.  The dynamic-instability machinery of
the reference driver is not part of this path.

Two modes:
  'plain'  : FP64 pair sums, FP64 fiber LU, zero initial guess (the reference
             method on the GPU);
  'fast'   : FP32-compensated pair sums inside the Krylov iteration with the
             FP64 operator deciding convergence, FP32 fiber LU with FP64
             refinement, and an optional initial-guess predictor.

Fibers may be clamped to rigid bodies (attached=[(fiber, body, link)],
tstep.py:878-881 with tstep_utils.build_link_matrix :948-1091): their minus
end follows the body and its tangent turns with it, and their end force and
moment act on the body.  Fibers listed in hinged hold their plus end in
place, free to rotate, as a fiber that reached the periphery
(tstep.py:886-888); cortex=True clamps the minus end of every other fiber in
place (attached_to_cortex, tstep.py:881-883).

Time steps are fixed by default.  adaptive=True follows the reference's
accept/reject scheme (tstep.py:427-581, check_error :1874-1925): after each
solve the fibers' inextensibility error max |(|x_s| - 1)| of the new
configuration decides; a step within tol_tstep is accepted (and the next dt
grows by 1.2, up to dt_max, when the error is below 0.9 tol_tstep), a step
above it, or one that puts a fiber node on or outside the periphery (a sphere
of radius periphery_radius, or an ellipsoid of semi-axes periphery_axes), is
rejected and retried with dt / 2 (down to dt_min, where it is accepted).

In fast mode the closing residual is, by default (certify="update"), the
guess's FP64 residual updated by the correction's image under the
FP32-compensated operator, accepted only under a bound that includes that
operator's error (fiber3d_amd.gmres), an error measured on the system's own
corrections; certify="fp64" recomputes b - A x in FP64
instead.  With record=True every step keeps its geometry, right-hand
side and solution, and verify() re-assembles each recorded step and returns
the true FP64 residuals.
"""
import time
from dataclasses import dataclass

import numpy as np
import torch

from .fileio import (fibers_from_file, write_body_velocities_frame, write_clones_frame,
                     write_fibers_frame, write_resume_clones, write_resume_fibers)
from .gmres import GmresContext, gmres
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
    accepted: bool = True
    dt: float = 0.0                  # the step size this solve used
    inextensibility: float = 0.0     # max | |x_s| - 1 | of the new configuration
    guess_residual: float = None     # ||b - A x0|| / ||b|| of an accepted initial guess


def periphery_collision(X, a, b, c):
    """True if a fiber node lies on or outside the ellipsoidal periphery of
    semi-axes (a, b, c) (the fiber-cortex part of tstep.check_collision,
    tstep.py:1945-1987, with its arithmetic: the node's angles in the
    axis-scaled frame give a periphery point, and the node collides when it
    is at least that point's distance from the origin).  X: (n, 3) tensor."""
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    xs, ys, zs = x / a, y / b, z / c
    r_true = torch.sqrt(x * x + y * y + z * z)
    r_scaled = torch.sqrt(xs * xs + ys * ys + zs * zs)
    phi = torch.atan2(ys, xs + 1e-12)
    theta = torch.arccos(zs / (1e-12 + r_scaled))
    px = a * torch.sin(theta) * torch.cos(phi)
    py = b * torch.sin(theta) * torch.sin(phi)
    pz = c * torch.cos(theta)
    return bool((r_true >= torch.sqrt(px * px + py * py + pz * pz)).any())


def fibers_reaching_periphery(fibers, a, b, c, margin=0.05):
    """Indices of the fibers that reach the periphery of semi-axes (a, b, c),
    by the reference's test (tstep_utils.py:343-361): some node is on or
    outside it, or within margin times the periphery point's distance of
    that point (the point at the node's angles in the axis-scaled frame).
    The reference hinges such fibers' plus ends; pass the result as hinged.
    fibers: list of (N_k, 3) host arrays."""
    out = []
    for k, x in enumerate(fibers):
        x = np.asarray(x, dtype=np.float64)
        xs, ys, zs = x[:, 0] / a, x[:, 1] / b, x[:, 2] / c
        r_true = np.sqrt((x * x).sum(1))
        r_scaled = np.sqrt(xs * xs + ys * ys + zs * zs)
        phi = np.arctan2(ys, xs + 1e-12)
        theta = np.arccos(zs / (1e-12 + r_scaled))
        p = np.stack([a * np.sin(theta) * np.cos(phi), b * np.sin(theta) * np.sin(phi),
                      c * np.cos(theta)], axis=1)
        d2p = np.sqrt(((x - p) ** 2).sum(1))
        pr = np.sqrt((p * p).sum(1))
        if (r_true >= pr).any() or (d2p <= margin * pr).any():
            out.append(k)
    return out


def adaptive_decision(err, collision, dt, tol_tstep, dt_min, dt_max):
    """(accept, next dt) of tstep.check_error (tstep.py:1900-1922)."""
    if err <= tol_tstep:
        accept = True
        dt_new = min(dt_max, dt * 1.2) if err <= 0.9 * tol_tstep else dt
    else:
        accept, dt_new = False, dt * 0.5
    if collision:
        accept, dt_new = False, dt * 0.5
    if dt_new < dt_min:
        accept, dt_new = True, dt_min
    return accept, dt_new


class TimeStepper:
    def __init__(self, fibers, shell, mode="plain", eta=1.0, dt=1e-3, E=0.01, epsilon=1e-3,
                 penalty=500.0, nfd=4, end_force=1.0, discretisation="reference", tol=1e-8,
                 restart=150, maxiter=200, predictor=None, lengths=None, bodies=None,
                 body_force=None, graph=False, certify="update", record=False, lu_every=None,
                 target=None, adaptive=False, tol_tstep=1e-2, dt_min=1e-5, dt_max=1e-2,
                 periphery_radius=None, attached=None, body_orientations=None, hinged=None,
                 cortex=False, periphery_axes=None):
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
        # Relative error of the FP32-compensated operator on a step's
        # correction, which the updated-residual certificate adds to its
        # bound.  It depends on the geometry and the state (targets close to
        # a surface, large cancelling forces of a fiber held at both ends),
        # so it is measured on the corrections of guess-started solves,
        # free of cost whenever a solve closes with the FP64 residual.  When
        # no solve has for 2 lu_every steps, and after every re-factorisation,
        # change of dt or rejected step, the next solve that iterates is
        # closed in FP64 to measure it (one FP64 product).  The bound uses
        # the largest ratio of the last 4 lu_every steps times a safety
        # factor, never below the floor; until the first measurement every
        # residual is recomputed in FP64.
        self.inner_rel_err = None
        self.inner_rel_err_floor = 1e-6
        self.inner_rel_err_safety = 10.0
        self._inner_meas = []    # (step index, measured ratio)
        self._changed = False    # the last assembly re-factored or changed dt
        self.adaptive = adaptive
        self.tol_tstep, self.dt_min, self.dt_max = tol_tstep, dt_min, dt_max
        self.periphery_radius = periphery_radius
        self.periphery_axes = None if periphery_axes is None else tuple(map(float, periphery_axes))
        self.time = 0.0          # time of the last accepted solution
        self.rejected = 0
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
                               lu_refine=0, attached=attached,
                               body_orientations=body_orientations, hinged=hinged,
                               cortex=cortex)
        self.predictor = predictor
        S = self.sys
        self.graphs = bool(graph)
        self.gctx = None
        self._last_its = 0
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
        self.body_velocity = torch.zeros((S.nbody, 6), dtype=torch.float64, device=dev)
        self.history = []

    def inner_operator(self):
        S = self.sys
        if self.mode == "fast":
            return lambda v, out=None: S.matvec(v, out=out, precision="f32c")
        return None

    def _note_inner(self, e):
        self._inner_meas.append((self.nsteps, float(e)))
        self._inner_meas = [(k, v) for k, v in self._inner_meas
                            if k > self.nsteps - 4 * self.lu_every]
        worst = max(v for _, v in self._inner_meas)
        self.inner_rel_err = max(self.inner_rel_err_floor, self.inner_rel_err_safety * worst)

    def inextensibility_error(self, as_tensor=False):
        """max over all fiber nodes of | |x_s| - 1 | for the assembled
        configuration (tstep.check_error, tstep.py:1874-1891); as a device
        scalar when as_tensor is set (no host synchronisation)."""
        e = (torch.linalg.norm(self.sys.xs, dim=1) - 1.0).abs().max()
        return e if as_tensor else float(e)

    def collision(self):
        """A fiber node on or outside the periphery (the fiber-cortex part of
        tstep.check_collision, tstep.py:1927-1987); no periphery, no
        collision."""
        if self.periphery_axes is not None:
            return periphery_collision(self.sys.X, *self.periphery_axes)
        if self.periphery_radius is None:
            return False
        R = self.periphery_radius
        return periphery_collision(self.sys.X, R, R, R)

    def step(self):
        S = self.sys
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dt = S.dt
        x0 = None
        if self.predictor is not None:
            x0 = self.predictor.predict(self)
        # device GMRES: buffers (basis, state) reused by every step; with
        # graph=True each Arnoldi step (basis vector -> A P^-1 -> CGS2 ->
        # Givens) replays one recorded HIP graph per operator precision
        if self.gctx is None:
            self.gctx = GmresContext(S.size, min(self.restart, S.size), S.dev)
        fused = {"ctx": self.gctx, "expect": max(0, self._last_its - 1)}
        if self.graphs:
            fused["graph_keys"] = {"matvec": "f64", "inner": "f32c"}
        certify_update = self.mode == "fast" and self.certify == "update"
        # after a re-factorisation, a change of dt or a rejected step the
        # measured error may no longer describe the operator: close in FP64
        measure_due = (not self._inner_meas or self._changed
                       or self.nsteps - self._inner_meas[-1][0] >= 2 * self.lu_every)
        res = gmres(S.matvec, S.rhs, S.precond, tol=self.tol, restart=self.restart,
                    maxiter=self.maxiter, x0=x0, inner=self.inner_operator(),
                    update_residual=certify_update and not measure_due,
                    inner_rel_err=self.inner_rel_err or 1e-6, target=self.target,
                    measure_inner=certify_update, **fused)
        if certify_update and res.guess_used and res.inner_err is not None:
            self._note_inner(res.inner_err)
        self._last_its = res.iterations
        sol = res.x
        # the residual correction P^-1 r for the predictor, with this step's
        # preconditioner (before the re-assembly below)
        corr = gain = None
        if (self.predictor is not None and getattr(self.predictor, "wants_correction", False)
                and res.resid is not None):
            corr = S.precond(res.resid)
            if getattr(self.predictor, "needs_gain", False):
                # |r - A P^-1 r| / |r| with this step's operator (FP64)
                rr = res.resid
                gain = float(torch.linalg.norm(rr - S.matvec(corr)) /
                             torch.linalg.norm(rr).clamp_min(1e-300))
        body_old = S.bodies.state() if S.bodies is not None else None
        if self.record:
            # S.rhs and sol are fresh tensors of this step; X is updated in place
            self.records.append((S.X.clone(), body_old, S.rhs, sol, dt))
        X_old = S.X.clone() if self.adaptive else None
        S.X.copy_(sol[self.idx_x])
        if S.bodies is not None:
            # rigid motion with the solved velocities (tstep.py:1398-1406)
            S.bodies.move(S.bodies._vel(sol, 3 * S.nper), dt)
            # attached fibers' minus ends back onto their sites (tstep.py:462-470)
            S.fix_links()
        # re-assembly refills the system's buffers in place: the recorded
        # Arnoldi graphs stay valid
        self.nsteps += 1
        refactor = self.nsteps % self.lu_every == 0
        S.prepare(refactor=refactor)
        self._changed = refactor
        # read at once only when the step size depends on it, else after
        # the step's closing synchronisation
        err = self.inextensibility_error(as_tensor=not self.adaptive)
        accept, dt_new = True, dt
        if self.adaptive:
            accept, dt_new = adaptive_decision(err, self.collision(), dt, self.tol_tstep,
                                               self.dt_min, self.dt_max)
        if accept:
            self.time += dt
            if self.predictor is not None:
                if getattr(self.predictor, "wants_correction", False):
                    if gain is not None:
                        self.predictor.observe(self, sol, corr, gain)
                    else:
                        self.predictor.observe(self, sol, corr)
                else:
                    self.predictor.observe(self, sol)
            self.prev_solution = sol
            self.shell_density = sol[:3 * S.nper].clone()
            self.tension = sol[self.idx_t]
            if S.bodies is not None:
                self.body_velocity = S.bodies._vel(sol, 3 * S.nper).clone()
        else:
            # rejected (tstep.py:540-579): back to the configuration it left
            self.rejected += 1
            S.X.copy_(X_old)
            if S.bodies is not None:
                S.bodies.set_state(body_old)
        if not accept or dt_new != dt:
            S.dt = dt_new
            S.prepare(refactor=True)
            self._changed = True
        torch.cuda.synchronize()
        err = float(err)
        st = StepStats(res.iterations, res.restarts, res.residual, time.perf_counter() - t0,
                       res.guess_used, res.residual_kind, res.residual_bound, accept, dt, err,
                       res.history[0] if res.guess_used and res.history else None)
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

    def write_bodies(self, f_clones, f_velocities=None, radius=None, txt=True):
        """Append the bodies' locations and orientations (and, with
        f_velocities, their last velocities) as one frame of the reference's
        clones output (tstep.py:2088-2121).  radius defaults to each body's
        mean node distance from its centre."""
        B = self.sys.bodies
        if radius is None:
            r = torch.linalg.norm(B.pos - B.centres[B.bid], dim=1)
            radius = torch.zeros(B.count, dtype=torch.float64, device=r.device) \
                .index_add_(0, B.bid, r).cpu().numpy() / B.nb
        write_clones_frame(f_clones, B.centres.cpu().numpy(), B.quat.cpu().numpy(), radius,
                           txt=txt)
        if f_velocities is not None:
            write_body_velocities_frame(f_velocities, self.body_velocity.cpu().numpy())

    def write_resume_bodies(self, name_file):
        """The reference's <output>_<ID>_resume.clones for the current bodies."""
        B = self.sys.bodies
        write_resume_clones(name_file, B.centres.cpu().numpy(), B.quat.cpu().numpy())

    def verify(self):
        """True relative residuals ||b - A x|| / ||b|| of the recorded steps in
        FP64, each with the system re-assembled at that step's geometry, and
        the largest difference between the re-assembled and the recorded
        right-hand sides.  The current state is restored afterwards."""
        S = self.sys
        now = (S.X.clone(), S.bodies.state() if S.bodies is not None else None, S.dt)
        out, drhs = [], 0.0
        for X, body, b, x, dt in self.records:
            S.X.copy_(X)
            if body is not None:
                S.bodies.set_state(body)
            S.dt = dt
            S.prepare()
            drhs = max(drhs, float((S.rhs - b).abs().max()))
            r = b - S.matvec(x, precision="f64")
            out.append(float(torch.linalg.norm(r) / torch.linalg.norm(b)))
        S.X.copy_(now[0])
        if now[1] is not None:
            S.bodies.set_state(now[1])
        S.dt = now[2]
        S.prepare()
        return out, drhs


def stepper_from_input(read, shell, **kw):
    """A TimeStepper set up from a ReadInput (read_input/read_input.py) and a
    Shell: the fibers of every structure_fibers file (each fiber's E and
    length from its header, else E and length_fibers), dt, eta, epsilon,
    nfd = num_points_finite_diff (4 when unset) and the GMRES tolerance
    solver_tolerance.  Keyword arguments are passed on and take precedence.
    One elastic modulus serves all fibers, so differing header values are
    rejected."""
    fibers, Es, Ls = [], [], []
    for name in read.structures_fibers:
        x, E, L = fibers_from_file(name, E=read.E, length=read.length_fibers)
        fibers += x
        Es += E
        Ls += L
    if not fibers:
        raise ValueError("the input file names no structure_fibers file")
    bad = [k for k, L in enumerate(Ls) if not L > 0.0]
    if bad:
        raise ValueError(f"fibers {bad[:8]} have no positive length: give it in the fiber "
                         "file header or as length_fibers in the input file")
    if len(set(Es)) != 1:
        raise ValueError("fibers with different E are not supported: %s" % sorted(set(Es)))
    if not read.num_points_finite_diff and "nfd" not in kw:
        # the reference passes the option unchanged to fiber_matrices
        # (tstep.py:262), where 0 builds one-point stencils (D_1 = 0)
        raise ValueError("num_points_finite_diff is unset (0) in the input file: set it, or "
                         "pass nfd=")
    args = dict(dt=read.dt, eta=read.eta, E=Es[0], epsilon=read.epsilon,
                nfd=read.num_points_finite_diff, tol=read.solver_tolerance,
                lengths=np.asarray(Ls, dtype=np.float64))
    args.update(kw)
    return TimeStepper(fibers, shell, **args)
