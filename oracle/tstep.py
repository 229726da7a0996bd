"""Oracle: one implicit hydrodynamic time step of fibers in a periphery (FP64).

Restates, for free fibers (force/torque-free ends) and free rigid bodies (no
fiber attached) inside a spherical shell, the path of tstep.time_step_hydro
(python/tstep/tstep.py:628-1216):
  * explicit flows of the external fiber forces and the system RHS
    (tstep.py:757-846, tstep_utils.get_fibers_and_bodies_matrices
    tstep_utils.py:1238-1336, flow_fibers :653-700, self_flow_fibers :812-851);
  * the matvec A_body_fiber_hyd (tstep.py:1003-1178);
  * the block preconditioner P_inv (tstep.py:950-990);
  * right-preconditioned restarted GMRES (bpm_utilities/gmres.py:94-108 over
    SciPy's gmres, tstep.py:1209-1216), here restated as textbook GMRES
    (Arnoldi with modified Gram-Schmidt, Givens rotations).
  * rigid bodies: surface double layer with singularity subtraction, K = [J, rot]
    (body/body.py:162-193, :325-346), the body rows of the matvec
    (tstep.py:1044-1143), external force/torque as a Stokeslet and rotlet at
    the centre (tstep.py:818-846), the dense body preconditioner
    (tstep_utils.py:1341-1387).
The unknown vector is [shell density (3 Nper, AoS) | per body density (3 Nb,
AoS), U, Omega | per fiber X, Y, Z, T].

Pinned by tests/test_oracle_tstep.py: the matvec against an independently
assembled dense operator (fibers and shell), the preconditioner blocks against
P^-1 (A x) = x, GMRES against a dense direct solve, and the body formulation
against Stokes' law for a lone sphere (U = F/(6 pi eta a), Omega =
L/(8 pi eta a^3)).  The reference's matvec
and SciPy's GMRES cannot be imported here (stkfmm / mpi4py / numba.errors
imports), so beyond those invariants the system assembly is parity unpinned.
"""
import numpy as np
import scipy.linalg as sla

from . import fiber as OF
from . import kernels as K
from . import periphery as P


class HydroSystem:
    def __init__(self, fibers, shell=None, eta=1.0, dt=1e-3, E=0.01, epsilon=1e-3,
                 penalty=500.0, nfd=4, end_force=1.0, discretisation="reference",
                 lengths=None, bodies=None, body_force=None, attached=None, hinged=None,
                 cortex=False, shell_cache=None):
        """fibers: list of (N_k,3) arrays; shell: (x, normals, weights) or None;
        bodies: list of (centre (3,), nodes (Nb,3), outward normals, weights);
        body_force: optional (nbody, 6) external force and torque per body;
        attached: optional list of (fiber k, body j, link (3,)): fiber k's
        minus end is clamped to body j at centre_j + link (tstep.py:878-881,
        tstep_utils.build_link_matrix :948-1091); hinged: fibers whose plus
        end is held in place with free rotation, as a fiber that reached the
        periphery (tstep.py:886-888); cortex: every fiber not attached to a
        body has its minus end clamped in place (attached_to_cortex,
        tstep.py:881-883).  shell_cache: the `shell_state()` of an earlier
        system with the same shell, whose one-time set-up (subtraction
        vectors, dense shell matrix and its LU, tstep.py:278-356) is reused."""
        self.eta, self.dt, self.E, self.penalty = eta, dt, E, penalty
        self.c0, self.c1 = OF.constants(epsilon, eta)
        self.fibers = [np.asarray(x, float) for x in fibers]
        self.sizes = np.array([x.shape[0] for x in self.fibers])
        self.offs = np.concatenate([[0], np.cumsum(self.sizes)])
        self.lengths = np.full(len(fibers), 1.0) if lengths is None else np.asarray(lengths, float)
        self.fm = {N: OF.FiberMatrices(N, nfd, discretisation) for N in set(self.sizes.tolist())}
        self.npts = int(self.offs[-1])
        self.shell = shell
        self.nper = 0 if shell is None else shell[0].shape[0]
        self.bodies = [] if bodies is None else [tuple(np.asarray(a, float) for a in b) for b in bodies]
        self.nbody = len(self.bodies)
        self.bsizes = np.array([b[1].shape[0] for b in self.bodies], dtype=int)
        self.boffs = np.concatenate([[0], np.cumsum(self.bsizes)]).astype(int)
        self.body_force = np.zeros((self.nbody, 6)) if body_force is None else np.asarray(body_force, float)
        # unknowns: shell | per body (3 Nb density, 6 velocity) | fibers
        self.fib0 = 3 * self.nper + 3 * int(self.boffs[-1]) + 6 * self.nbody
        self.size = self.fib0 + 4 * self.npts
        self.trg_fib = np.concatenate(self.fibers) if self.fibers else np.zeros((0, 3))
        self.attached = [] if attached is None else [(int(k), int(j), np.asarray(l, float))
                                                     for k, j, l in attached]
        self.clamped = {k for k, _, _ in self.attached}
        self.cortex = sorted(set(range(len(self.fibers))) - self.clamped) if cortex else []
        self.clamped |= set(self.cortex)
        self.hinged = set() if hinged is None else {int(k) for k in hinged}
        self._shell_cache = shell_cache
        self._setup(end_force)

    def shell_state(self):
        """The shell's one-time set-up, for shell_cache."""
        return {"e": (self.ex, self.ey, self.ez), "M": self.M_shell, "lu": self.lu_shell}

    # -------------------------------------------------------------------
    def _setup(self, end_force):
        eta = self.eta
        # per-fiber derivatives and matrices (tstep.py:689-718)
        self.derivs = []
        self.weights = []
        self.F_blocks = []
        self.D1 = []
        for k, x in enumerate(self.fibers):
            fm = self.fm[x.shape[0]]
            L = self.lengths[k]
            self.derivs.append(OF.derivatives(fm, x, L))
            self.weights.append(fm.weights(L))
            self.F_blocks.append(OF.force_operator(fm, x, L, self.E))
            self.D1.append(fm.Ds(L)[0])
        # explicit force: plus end (last node) along the end tangent
        force = np.zeros((self.npts, 3))
        for k in range(len(self.fibers)):
            xs = self.derivs[k][0]
            force[self.offs[k + 1] - 1] = end_force * xs[-1]
        self.force = force
        # explicit flows (tstep.py:794-846): fibers -> fibers (self block
        # removed) and fibers -> shell
        fw = force * np.concatenate(self.weights)[:, None] if self.npts else np.zeros((0, 3))
        vfib2fib = self._fib2fib(fw)
        rhs = np.zeros(self.size)
        # bodies: surface geometry, subtraction vectors, K
        self.bnodes = np.concatenate([b[1] for b in self.bodies]) if self.nbody else np.zeros((0, 3))
        self.bnrm = np.concatenate([b[2] for b in self.bodies]) if self.nbody else np.zeros((0, 3))
        self.bw = np.concatenate([b[3] for b in self.bodies]) if self.nbody else np.zeros(0)
        self.be = []
        for (c, xb, nb, wb) in self.bodies:
            self.be.append(P.singularity_subtraction_vectors(xb, nb, wb, eta))
        # explicit flows of the external body force/torque: Stokeslet + rotlet at
        # each centre (tstep.py:818-846), on fibers, bodies and shell
        centres = np.array([b[0] for b in self.bodies]).reshape(-1, 3)
        def body_explicit(trg):
            if self.nbody == 0 or trg.shape[0] == 0:
                return np.zeros(3 * trg.shape[0])
            return (K.oseen_kernel_source_target(centres, trg, self.body_force[:, :3], eta)
                    + K.rotlet_kernel_source_target(centres, trg, self.body_force[:, 3:], eta))
        vfib2fib = vfib2fib + body_explicit(self.trg_fib).reshape(-1, 3)
        if self.nbody:
            vb = K.oseen_kernel_source_target(self.trg_fib, self.bnodes, fw, eta) \
                if self.npts else np.zeros(3 * self.bnodes.shape[0])
            vb = vb + body_explicit(self.bnodes)
            for k in range(self.nbody):
                o = self._body_off(k)
                rhs[o:o + 3 * self.bsizes[k]] = -vb[3 * self.boffs[k]:3 * self.boffs[k + 1]]
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            vfib2shell = K.oseen_kernel_source_target(self.trg_fib, xs_, fw, eta) \
                if self.npts else np.zeros(3 * self.nper)
            rhs[:3 * self.nper] = -(vfib2shell + body_explicit(xs_))
            if self._shell_cache is not None:
                self.ex, self.ey, self.ez = self._shell_cache["e"]
            else:
                self.ex, self.ey, self.ez = P.singularity_subtraction_vectors(xs_, ns_, ws_, eta)
        # per-fiber square systems (get_fibers_and_bodies_matrices)
        self.A_sq = []
        self.Q = []
        for k, x in enumerate(self.fibers):
            N = x.shape[0]
            fm = self.fm[N]
            L = self.lengths[k]
            f_k = force[self.offs[k]:self.offs[k + 1]]
            v_k = vfib2fib[self.offs[k]:self.offs[k + 1]]
            A = OF.form_linear_operator(fm, x, L, self.E, self.dt, self.c0, self.c1, self.penalty)
            R = OF.compute_RHS(fm, x, L, self.dt, self.c0, self.c1, self.penalty, force=f_k,
                               flow=v_k)
            A_sq, b = OF.apply_BC_rectangular(fm, x, L, self.E, A, R, end_force=f_k[-1],
                                              clamped=k in self.clamped, dt=self.dt,
                                              c0=self.c0, flow=v_k, force=f_k,
                                              hinged=k in self.hinged)
            self.A_sq.append(A_sq)
            o = self.fib0 + 4 * self.offs[k]
            rhs[o:o + 4 * N] = b
            # v_on_fib -> P_cheb [v; xs . D1 v], padded with 14 zero rows
            # (tstep.py:1150-1171)
            xs = self.derivs[k][0]
            T = np.zeros((4 * N, 3 * N))
            for c in range(3):
                T[c * N:(c + 1) * N, c * N:(c + 1) * N] = np.eye(N)
                T[3 * N:, c * N:(c + 1) * N] = xs[:, c][:, None] * self.D1[k]
            Q = np.zeros((4 * N, 3 * N))
            Q[:4 * N - 14] = fm.P_cheb @ T
            self.Q.append(Q)
        self.rhs = rhs
        # preconditioner factors (tstep.py:936-941, :348-351)
        self.lu_fib = [sla.lu_factor(A, check_finite=False) for A in self.A_sq]
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            if self._shell_cache is not None:
                self.M_shell, self.lu_shell = self._shell_cache["M"], self._shell_cache["lu"]
            else:
                self.M_shell = P.shell_matrix(xs_, ns_, ws_, self.ex, self.ey, self.ez, eta)
                self.lu_shell = sla.lu_factor(self.M_shell)
        # body preconditioner blocks [M - I_sub, -K; -K^T, I] (tstep_utils.py:1341-1387)
        self.K = []
        self.lu_body = []
        for k, (c, xb, nb, wb) in enumerate(self.bodies):
            Kb = self.k_matrix(xb, c)
            self.K.append(Kb)
            M = K.stresslet_kernel_times_normal(xb, nb, eta)
            ex, ey, ez = self.be[k]
            n = xb.shape[0]
            for i in range(n):
                M[3 * i:3 * i + 3, 3 * i + 0] -= ex[i] / wb[i]
                M[3 * i:3 * i + 3, 3 * i + 1] -= ey[i] / wb[i]
                M[3 * i:3 * i + 3, 3 * i + 2] -= ez[i] / wb[i]
            A = np.zeros((3 * n + 6, 3 * n + 6))
            A[:3 * n, :3 * n] = M
            A[:3 * n, 3 * n:] = -Kb
            A[3 * n:, :3 * n] = -Kb.T
            A[3 * n:, 3 * n:] = np.eye(6)
            self.lu_body.append(sla.lu_factor(A))

    @staticmethod
    def k_matrix(xb, centre):
        """K = [J, rot] (3Nb x 6): K [U; w] = U + w x (r - centre) (body.py:162-193)."""
        r = xb - centre
        n = r.shape[0]
        Kb = np.zeros((3 * n, 6))
        for i in range(n):
            Kb[3 * i:3 * i + 3, :3] = np.eye(3)
            x, y, z = r[i]
            Kb[3 * i:3 * i + 3, 3:] = [[0.0, z, -y], [-z, 0.0, x], [y, -x, 0.0]]
        return Kb

    def _body_off(self, k):
        return 3 * self.nper + 3 * int(self.boffs[k]) + 6 * k

    def _fib2fib(self, fw):
        """Oseen flow of fiber forces on fiber nodes minus each fiber's self block
        (flow_fibers + self_flow_fibers, tstep_utils.py:653-729, :812-851)."""
        if self.npts == 0:
            return np.zeros((0, 3))
        v = K.oseen_kernel_source_target(self.trg_fib, self.trg_fib, fw, self.eta).reshape(-1, 3)
        for k, x in enumerate(self.fibers):
            o0, o1 = self.offs[k], self.offs[k + 1]
            G = K.oseen_tensor(x, self.eta)
            v[o0:o1] -= (G @ fw[o0:o1].ravel()).reshape(-1, 3)
        return v

    # -------------------------------------------------------------------
    def split(self, x):
        q = x[:3 * self.nper].reshape(-1, 3)
        XT = x[self.fib0:]
        return q, XT

    def body_parts(self, x):
        lam, vel = [], []
        for k in range(self.nbody):
            o = self._body_off(k)
            n = self.bsizes[k]
            lam.append(x[o:o + 3 * n].reshape(-1, 3))
            vel.append(x[o + 3 * n:o + 3 * n + 6])
        return lam, vel

    def _link_forces(self, XT):
        """Force and torque of every attached fiber on its body (the body rows
        of tstep_utils.build_link_matrix, tstep_utils.py:995-1036):
        F = -(E X_sss(0) - T(0) x_s(0)), L = link x F + E x_s(0) x X_ss(0),
        summed per body -> (nbody, 6)."""
        out = np.zeros((self.nbody, 6))
        for k, j, link in self.attached:
            N = self.sizes[k]
            o = 4 * self.offs[k]
            X = XT[o:o + 3 * N].reshape(3, N)
            T0 = XT[o + 3 * N]
            _, D2, D3, _ = self.fm[N].Ds(self.lengths[k])
            xs0 = self.derivs[k][0][0]
            F = -self.E * (X @ D3[0]) + T0 * xs0
            out[j, :3] += F
            out[j, 3:] += np.cross(link, F) + self.E * np.cross(xs0, X @ D2[0])
        return out

    def matvec(self, x):
        """A_body_fiber_hyd (tstep.py:1003-1178) for fibers, free bodies and shell."""
        q, XT = self.split(x)
        lam, vel = self.body_parts(x)
        eta = self.eta
        res = np.zeros_like(x)
        centres = np.array([b[0] for b in self.bodies]).reshape(-1, 3)
        link_ft = self._link_forces(XT) if self.attached else None

        def link_flow(trg):
            """Stokeslet + rotlet of the fibers' forces and torques on the
            bodies, at the body centres (tstep.py:1052-1064)."""
            if link_ft is None or trg.shape[0] == 0:
                return 0.0
            return (K.oseen_kernel_source_target(centres, trg, link_ft[:, :3], eta)
                    + K.rotlet_kernel_source_target(centres, trg, link_ft[:, 3:], eta))
        f = np.zeros((self.npts, 3))
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            o = 4 * self.offs[k]
            fk = self.F_blocks[k] @ XT[o:o + 4 * N]
            f[self.offs[k]:self.offs[k + 1]] = fk.reshape(3, N).T
        fw = f * np.concatenate(self.weights)[:, None] if self.npts else f
        v_fib = self._fib2fib(fw)
        lam_all = np.concatenate(lam) if self.nbody else np.zeros((0, 3))
        if self.nbody:
            # body double layer on every target (the j = i term is exactly zero)
            if self.npts:
                v_fib = v_fib + K.stresslet_kernel_source_target(
                    self.bnodes, self.trg_fib, self.bnrm, lam_all, eta).reshape(-1, 3)
                v_fib = v_fib + np.reshape(link_flow(self.trg_fib), (-1, 3)) \
                    if link_ft is not None else v_fib
            vb = K.stresslet_kernel_source_target(self.bnodes, self.bnodes, self.bnrm, lam_all, eta)
            if self.npts:
                vb = vb + K.oseen_kernel_source_target(self.trg_fib, self.bnodes, fw, eta)
            vb = vb + link_flow(self.bnodes)
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            sh = self.M_shell @ q.ravel()
            if self.npts:
                sh = sh + K.oseen_kernel_source_target(self.trg_fib, xs_, fw, eta)
                v_fib = v_fib + K.stresslet_kernel_source_target(xs_, self.trg_fib, ns_, q,
                                                                 eta).reshape(-1, 3)
            if self.nbody:
                sh = sh + K.stresslet_kernel_source_target(self.bnodes, xs_, self.bnrm, lam_all, eta)
                vb = vb + K.stresslet_kernel_source_target(xs_, self.bnodes, ns_, q, eta)
                sh = sh + link_flow(xs_)
            res[:3 * self.nper] = sh
        for k in range(self.nbody):
            o = self._body_off(k)
            n = self.bsizes[k]
            wb = self.bodies[k][3]
            ex, ey, ez = self.be[k]
            l = lam[k]
            sub = (l[:, 0] / wb)[:, None] * ex + (l[:, 1] / wb)[:, None] * ey \
                + (l[:, 2] / wb)[:, None] * ez
            res[o:o + 3 * n] = vb[3 * self.boffs[k]:3 * self.boffs[k + 1]] - sub.ravel() \
                - self.K[k] @ vel[k]
            res[o + 3 * n:o + 3 * n + 6] = -self.K[k].T @ l.ravel() + vel[k]
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            o = 4 * self.offs[k]
            vk = v_fib[self.offs[k]:self.offs[k + 1]].T.ravel()
            res[self.fib0 + o:self.fib0 + o + 4 * N] = \
                self.A_sq[k] @ XT[o:o + 4 * N] - self.Q[k] @ vk
        # clamped minus ends: the fiber rows of the link matrix
        # (tstep_utils.py:1039-1088) and the flow term of the tension condition
        # (tstep.py:1162-1166)
        for k, j, link in self.attached:
            N = self.sizes[k]
            r0 = self.fib0 + 4 * self.offs[k] + 4 * N - 14
            U, W = vel[j][:3], vel[j][3:]
            xs0 = self.derivs[k][0][0]
            vU = U + np.cross(W, link)
            res[r0:r0 + 3] -= vU
            res[r0 + 3] += -np.dot(vU, xs0) + np.dot(v_fib[self.offs[k]], xs0)
            res[r0 + 4:r0 + 7] -= np.cross(W, link / np.linalg.norm(link))
        # minus ends clamped to the cortex: only the flow term (tstep.py:1162-1166)
        for k in self.cortex:
            rt = self.fib0 + 4 * self.offs[k] + 4 * self.sizes[k] - 11
            res[rt] += np.dot(v_fib[self.offs[k]], self.derivs[k][0][0])
        # hinged plus ends: the flow term of their tension condition
        # (tstep.py:1167-1168)
        for k in self.hinged:
            N = self.sizes[k]
            rt = self.fib0 + 4 * self.offs[k] + 4 * N - 4
            res[rt] += np.dot(v_fib[self.offs[k + 1] - 1], self.derivs[k][0][-1])
        return res

    def precond(self, x):
        """P_inv (tstep.py:950-990): shell LU, body LU, per-fiber LU solves."""
        y = np.empty_like(x)
        if self.shell is not None:
            y[:3 * self.nper] = sla.lu_solve(self.lu_shell, x[:3 * self.nper])
        for k in range(self.nbody):
            o = self._body_off(k)
            m = 3 * self.bsizes[k] + 6
            y[o:o + m] = sla.lu_solve(self.lu_body[k], x[o:o + m])
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            o = self.fib0 + 4 * self.offs[k]
            y[o:o + 4 * N] = sla.lu_solve(self.lu_fib[k], x[o:o + 4 * N])
        return y

    def dense_operator(self):
        """The full system matrix assembled from its blocks (independent of matvec).
        Fibers and shell only."""
        assert self.nbody == 0, "dense_operator covers fibers and shell"
        n = self.size
        A = np.zeros((n, n))
        W = np.concatenate(self.weights)
        npts, nper = self.npts, self.nper
        # fiber force operator as one matrix: XT -> fw (AoS, npts x 3)
        Fbig = np.zeros((3 * npts, 4 * npts))
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            for c in range(3):
                rows = 3 * (self.offs[k] + np.arange(N)) + c
                Fbig[rows, 4 * self.offs[k]:4 * self.offs[k] + 4 * N] = \
                    self.F_blocks[k][c * N:(c + 1) * N] * self.weights[k][:, None]
        G = K.oseen_tensor(self.trg_fib, self.eta)
        for k in range(len(self.fibers)):
            s = slice(3 * self.offs[k], 3 * self.offs[k + 1])
            G[s, s] = 0.0
        Vff = G @ Fbig  # fiber velocities (AoS) from XT
        Qbig = np.zeros((4 * npts, 3 * npts))
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            for c in range(3):
                cols = 3 * (self.offs[k] + np.arange(N)) + c
                Qbig[4 * self.offs[k]:4 * self.offs[k] + 4 * N, cols] = self.Q[k][:, c * N:(c + 1) * N]
        Ablk = sla.block_diag(*self.A_sq)
        A[3 * nper:, 3 * nper:] = Ablk - Qbig @ Vff
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            # fiber -> shell Oseen, shell -> fiber stresslet
            d = xs_[:, None, :] - self.trg_fib[None, :, :]
            r = np.linalg.norm(d, axis=2)
            Gs = (np.eye(3)[None, None] / r[:, :, None, None]
                  + d[:, :, :, None] * d[:, :, None, :] / r[:, :, None, None] ** 3) / (8 * np.pi * self.eta)
            Gs = Gs.transpose(0, 2, 1, 3).reshape(3 * nper, 3 * npts)
            A[:3 * nper, 3 * nper:] = Gs @ Fbig
            A[:3 * nper, :3 * nper] = self.M_shell
            d2 = self.trg_fib[:, None, :] - xs_[None, :, :]
            r2 = np.linalg.norm(d2, axis=2)
            c = (-3.0 / (4 * np.pi * self.eta)) * np.einsum("ijk,jk->ij", d2, ns_) / r2**5
            S = c[:, :, None, None] * d2[:, :, :, None] * d2[:, :, None, :]
            S = S.transpose(0, 2, 1, 3).reshape(3 * npts, 3 * nper)
            A[3 * nper:, :3 * nper] = -Qbig @ S
        return A


def periphery_collision(X, a, b, c):
    """Fiber-cortex collision of tstep.check_collision (tstep.py:1945-1987):
    per node, the periphery point at the node's angles in the axis-scaled
    frame; a node at least as far from the origin as that point collides."""
    X = np.asarray(X, float)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    xs, ys, zs = x / a, y / b, z / c
    r_true = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    r_scaled = np.sqrt(xs ** 2 + ys ** 2 + zs ** 2)
    phi = np.arctan2(ys, xs + 1e-12)
    theta = np.arccos(zs / (1e-12 + r_scaled))
    p = np.stack([a * np.sin(theta) * np.cos(phi), b * np.sin(theta) * np.sin(phi),
                  c * np.cos(theta)], axis=1)
    return bool((r_true >= np.linalg.norm(p, axis=1)).any())


def gmres(matvec, b, precond=None, tol=1e-8, restart=150, maxiter=200, x0=None):
    """Right-preconditioned restarted GMRES: solve (A P^-1) y = b, x = P^-1 y.

    Stops when ||b - A x|| <= tol ||b|| (atol = 0), as bpm_utilities/gmres.py
    configures SciPy's gmres (tstep.py:1209-1216).  Returns (x, info, iters)."""
    n = b.size
    Pinv = precond if precond is not None else (lambda v: v)
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return np.zeros(n), 0, 0
    x = np.zeros(n) if x0 is None else np.array(x0, float)
    total = 0
    for _outer in range(maxiter):
        r = b - matvec(x)
        beta = np.linalg.norm(r)
        if beta <= tol * bnorm:
            return x, 0, total
        V = np.zeros((restart + 1, n))
        H = np.zeros((restart + 1, restart))
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        V[0] = r / beta
        j_done = 0
        for j in range(restart):
            w = matvec(Pinv(V[j]))
            for i in range(j + 1):
                H[i, j] = np.dot(V[i], w)
                w -= H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total += 1
            j_done = j + 1
            if abs(g[j + 1]) <= tol * bnorm:
                break
        yk = sla.solve_triangular(H[:j_done, :j_done], g[:j_done])
        x = x + Pinv(V[:j_done].T @ yk)
    r = b - matvec(x)
    return x, (0 if np.linalg.norm(r) <= tol * bnorm else total), total
