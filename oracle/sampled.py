"""Oracle: selected rows of the implicit fiber step at full problem size (FP64).

TEST INFRASTRUCTURE ONLY.  HydroSystem (oracle/tstep.py) assembles every fiber
block, the dense shell matrix and all O(N^2) sums in NumPy: fine up to a few
thousand points, hopeless at 262144 fiber nodes and a 32768-node periphery
(the dense shell operator alone is 77 GB).  SampledHydro restates the same
system -- tstep.time_step_hydro (python/tstep/tstep.py:628-1216) for free
fibers, free rigid bodies and a spherical periphery -- but evaluates only the
rows a parity test samples:

  * every fiber's O(N) quantities (tangents, quadrature weights, force
    operator, tstep.py:689-718, fiber/fiber.py:1399-1431) are formed, since
    every fiber is a source;
  * the square system A_sq and right-hand side b of the sampled fibers only
    (fiber.form_linear_operator / compute_RHS / apply_BC_rectangular,
    fiber/fiber.py:322-961, tstep_utils.py:1238-1336);
  * the O(N^2) sums only onto the sampled fiber nodes, the sampled shell
    nodes and the body nodes, with the C restatement of the reference's
    numba loops (oracle/nbody_ref.c, kernels/kernels.py:794-846, :1186-1217);
    the fibers' self blocks are added and then subtracted exactly as the
    reference does (tstep_utils.py:703-729, :812-851);
  * shell rows follow the dense operator of tstep.py:328-345 row by row:
    sum_{j != i} S(x_i - x_j; n_j) q_j - sum_c (q_ic / w_i) e_c,i - q_i / w_i
    + n_i sum_j n_j . q_j, with e_c (periphery.py:75-98) evaluated at the
    sampled nodes only.

Pinned by tests/test_oracle_sampled.py: at sizes HydroSystem handles, every
sampled row (A_sq, right-hand side, matvec) equals HydroSystem's.
"""
import numpy as np

from . import cbaseline as C
from . import fiber as OF
from . import kernels as K
from . import periphery as P


class SampledHydro:
    def __init__(self, fibers, shell=None, fiber_rows=(), shell_rows=(), eta=1.0, dt=1e-3,
                 E=0.01, epsilon=1e-3, penalty=500.0, nfd=4, end_force=1.0,
                 discretisation="reference", lengths=None, bodies=None, body_force=None,
                 reg=5e-3, eps=1e-5):
        """fibers: list of (N_k, 3) arrays; shell: (x, inward normals, weights) or
        None; bodies: list of (centre, nodes, outward normals, weights);
        fiber_rows: fiber indices whose rows are evaluated; shell_rows: shell
        node indices whose rows are evaluated (body rows are always all)."""
        self.eta, self.dt, self.E, self.penalty = eta, dt, E, penalty
        self.reg, self.eps = reg, eps
        self.c0, self.c1 = OF.constants(epsilon, eta)
        self.fibers = [np.asarray(x, float) for x in fibers]
        self.sizes = np.array([x.shape[0] for x in self.fibers], dtype=np.int64)
        self.offs = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.npts = int(self.offs[-1])
        self.lengths = np.full(len(fibers), 1.0) if lengths is None else np.asarray(lengths, float)
        self.fm = {N: OF.FiberMatrices(N, nfd, discretisation) for N in set(self.sizes.tolist())}
        self.shell = shell
        self.nper = 0 if shell is None else shell[0].shape[0]
        self.bodies = [] if bodies is None else [tuple(np.asarray(a, float) for a in b)
                                                 for b in bodies]
        self.nbody = len(self.bodies)
        self.bsizes = np.array([b[1].shape[0] for b in self.bodies], dtype=np.int64)
        self.boffs = np.concatenate([[0], np.cumsum(self.bsizes)]).astype(np.int64)
        self.body_force = np.zeros((self.nbody, 6)) if body_force is None \
            else np.asarray(body_force, float)
        self.fib0 = 3 * self.nper + 3 * int(self.boffs[-1]) + 6 * self.nbody
        self.size = self.fib0 + 4 * self.npts
        self.fiber_rows = [int(k) for k in fiber_rows]
        self.shell_rows = np.asarray(shell_rows, dtype=np.int64)
        self.trg_fib = np.concatenate(self.fibers)
        self._setup(end_force)

    # -- index sets of the sampled rows in the unknown vector ------------------
    def row_index(self):
        """Indices (into the full system vector) of the rows this oracle
        evaluates, in the order matvec/rhs return them: sampled shell nodes
        (3 per node, AoS), all body rows, sampled fibers (4N each)."""
        idx = [np.stack([3 * self.shell_rows + c for c in range(3)], axis=1).ravel()]
        idx.append(np.arange(3 * self.nper, self.fib0))
        for k in self.fiber_rows:
            o = self.fib0 + 4 * self.offs[k]
            idx.append(np.arange(o, o + 4 * self.sizes[k]))
        return np.concatenate(idx).astype(np.int64)

    def _sample_nodes(self):
        return np.concatenate([np.arange(self.offs[k], self.offs[k + 1])
                               for k in self.fiber_rows]).astype(np.int64) \
            if self.fiber_rows else np.zeros(0, np.int64)

    # --------------------------------------------------------------------------
    def _setup(self, end_force):
        eta = self.eta
        self.derivs, self.weights, self.F_blocks, self.D1 = [], [], [], []
        for k, x in enumerate(self.fibers):
            fm = self.fm[x.shape[0]]
            L = self.lengths[k]
            self.derivs.append(OF.derivatives(fm, x, L))
            self.weights.append(fm.weights(L))
            self.F_blocks.append(OF.force_operator(fm, x, L, self.E))
            self.D1.append(fm.Ds(L)[0])
        self.W = np.concatenate(self.weights)
        force = np.zeros((self.npts, 3))
        for k in range(len(self.fibers)):
            force[self.offs[k + 1] - 1] = end_force * self.derivs[k][0][-1]
        self.force = force
        fw = force * self.W[:, None]
        self.bnodes = np.concatenate([b[1] for b in self.bodies]) if self.nbody else np.zeros((0, 3))
        self.bnrm = np.concatenate([b[2] for b in self.bodies]) if self.nbody else np.zeros((0, 3))
        self.be = [P.singularity_subtraction_vectors(xb, nb, wb, eta)
                   for (c, xb, nb, wb) in self.bodies]
        self.K = [self._k_matrix(xb, c) for (c, xb, nb, wb) in self.bodies]
        sn = self._sample_nodes()
        # explicit flows at the sampled targets (tstep.py:794-846)
        vfib = self._fib_on(sn, fw) + self._body_explicit(self.trg_fib[sn])
        rhs = {}
        if self.nbody:
            vb = self._oseen(self.trg_fib, self.bnodes, fw) + self._body_explicit(self.bnodes)
            rhs["body"] = []
            for k in range(self.nbody):
                rhs["body"].append(np.concatenate(
                    [-vb[self.boffs[k]:self.boffs[k + 1]].ravel(), np.zeros(6)]))
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            xt = xs_[self.shell_rows]
            rhs["shell"] = -(self._oseen(self.trg_fib, xt, fw) + self._body_explicit(xt)).ravel()
            # subtraction vectors at the sampled nodes: e_c = S-self-sum of w e_c
            self.e_rows = []
            for c in range(3):
                dens = np.zeros((self.nper, 3))
                dens[:, c] = ws_
                self.e_rows.append(self._stresslet_self_rows(dens))
        # sampled fibers' square systems
        self.A_sq, self.b_fib, self.Q = {}, {}, {}
        pos = 0
        for k in self.fiber_rows:
            x = self.fibers[k]
            N = x.shape[0]
            fm = self.fm[N]
            L = self.lengths[k]
            f_k = force[self.offs[k]:self.offs[k + 1]]
            v_k = vfib[pos:pos + N]
            pos += N
            A = OF.form_linear_operator(fm, x, L, self.E, self.dt, self.c0, self.c1, self.penalty)
            R = OF.compute_RHS(fm, x, L, self.dt, self.c0, self.c1, self.penalty, force=f_k,
                               flow=v_k)
            A_sq, b = OF.apply_BC_rectangular(fm, x, L, self.E, A, R, end_force=f_k[-1])
            self.A_sq[k] = A_sq
            self.b_fib[k] = b
            xs = self.derivs[k][0]
            T = np.zeros((4 * N, 3 * N))
            for c in range(3):
                T[c * N:(c + 1) * N, c * N:(c + 1) * N] = np.eye(N)
                T[3 * N:, c * N:(c + 1) * N] = xs[:, c][:, None] * self.D1[k]
            Q = np.zeros((4 * N, 3 * N))
            Q[:4 * N - 14] = fm.P_cheb @ T
            self.Q[k] = Q
        self._rhs = rhs

    # -- pair sums (C restatement of the numba loops) ---------------------------
    def _oseen(self, src, trg, f):
        if src.shape[0] == 0 or trg.shape[0] == 0:
            return np.zeros((trg.shape[0], 3))
        return C.oseen(src, trg, f, self.eta, self.reg, self.eps).reshape(-1, 3)

    def _stresslet(self, src, trg, nrm, q):
        if src.shape[0] == 0 or trg.shape[0] == 0:
            return np.zeros((trg.shape[0], 3))
        return C.stresslet(src, trg, nrm, q, self.eta, self.reg, self.eps).reshape(-1, 3)

    def _stresslet_self_rows(self, q):
        """Shell self-sum (j != i) at the sampled shell nodes.  The i == j
        pair has d = 0 and contributes 0 in the C loop, as the reference's
        loop that skips it (kernels.py:1123-1183)."""
        xs_, ns_, _ = self.shell
        return self._stresslet(xs_, xs_[self.shell_rows], ns_, q)

    def _fib_on(self, nodes, fw):
        """Oseen flow of all fiber forces on the given fiber nodes, each node's
        own fiber's self block subtracted (flow_fibers + self_flow_fibers,
        tstep_utils.py:653-729, :812-851)."""
        v = self._oseen(self.trg_fib, self.trg_fib[nodes], fw)
        pos = 0
        for k in self.fiber_rows:
            o0, o1 = self.offs[k], self.offs[k + 1]
            G = K.oseen_tensor(self.fibers[k], self.eta, self.reg, self.eps)
            v[pos:pos + o1 - o0] -= (G @ fw[o0:o1].ravel()).reshape(-1, 3)
            pos += o1 - o0
        return v

    def _body_explicit(self, trg):
        """External body force/torque as Stokeslet + rotlet at the centres
        (tstep.py:818-846)."""
        if self.nbody == 0 or trg.shape[0] == 0 or not np.any(self.body_force):
            return np.zeros((trg.shape[0], 3))
        centres = np.array([b[0] for b in self.bodies])
        return (K.oseen_kernel_source_target(centres, trg, self.body_force[:, :3], self.eta)
                + K.rotlet_kernel_source_target(centres, trg, self.body_force[:, 3:],
                                                self.eta)).reshape(-1, 3)

    @staticmethod
    def _k_matrix(xb, centre):
        """K = [J, rot] (body.py:162-193)."""
        r = xb - centre
        n = r.shape[0]
        Kb = np.zeros((3 * n, 6))
        Kb[:, :3] = np.tile(np.eye(3), (n, 1))
        for i in range(n):
            x, y, z = r[i]
            Kb[3 * i:3 * i + 3, 3:] = [[0.0, z, -y], [-z, 0.0, x], [y, -x, 0.0]]
        return Kb

    def _body_off(self, k):
        return 3 * self.nper + 3 * int(self.boffs[k]) + 6 * k

    # --------------------------------------------------------------------------
    def rhs(self):
        """Right-hand side at row_index()."""
        parts = [self._rhs.get("shell", np.zeros(0))]
        parts += self._rhs.get("body", [])
        parts += [self.b_fib[k] for k in self.fiber_rows]
        return np.concatenate(parts)

    def matvec(self, x):
        """A_body_fiber_hyd (tstep.py:1003-1178) at row_index()."""
        eta = self.eta
        q = x[:3 * self.nper].reshape(-1, 3)
        XT = x[self.fib0:]
        lam, vel = [], []
        for k in range(self.nbody):
            o = self._body_off(k)
            n = self.bsizes[k]
            lam.append(x[o:o + 3 * n].reshape(-1, 3))
            vel.append(x[o + 3 * n:o + 3 * n + 6])
        lam_all = np.concatenate(lam) if self.nbody else np.zeros((0, 3))
        f = np.zeros((self.npts, 3))
        for k in range(len(self.fibers)):
            N = self.sizes[k]
            o = 4 * self.offs[k]
            f[self.offs[k]:self.offs[k + 1]] = (self.F_blocks[k] @ XT[o:o + 4 * N]).reshape(3, N).T
        fw = f * self.W[:, None]
        sn = self._sample_nodes()
        xt = self.trg_fib[sn]
        v_fib = self._fib_on(sn, fw)
        out = []
        if self.shell is not None:
            xs_, ns_, ws_ = self.shell
            r = self.shell_rows
            ps = xs_[r]
            sh = self._stresslet_self_rows(q)
            w = ws_[r]
            sub = sum((q[r, c] / w)[:, None] * self.e_rows[c] for c in range(3))
            sh = sh - sub - q[r] / w[:, None] + ns_[r] * np.sum(ns_ * q)
            sh = sh + self._oseen(self.trg_fib, ps, fw)
            if self.nbody:
                sh = sh + self._stresslet(self.bnodes, ps, self.bnrm, lam_all)
            out.append(sh.ravel())
            v_fib = v_fib + self._stresslet(xs_, xt, ns_, q)
        if self.nbody:
            v_fib = v_fib + self._stresslet(self.bnodes, xt, self.bnrm, lam_all)
            vb = self._stresslet(self.bnodes, self.bnodes, self.bnrm, lam_all) \
                + self._oseen(self.trg_fib, self.bnodes, fw)
            if self.shell is not None:
                vb = vb + self._stresslet(self.shell[0], self.bnodes, self.shell[1], q)
            for k in range(self.nbody):
                wb = self.bodies[k][3]
                ex, ey, ez = self.be[k]
                l = lam[k]
                sub = (l[:, 0] / wb)[:, None] * ex + (l[:, 1] / wb)[:, None] * ey \
                    + (l[:, 2] / wb)[:, None] * ez
                rb = vb[self.boffs[k]:self.boffs[k + 1]] - sub
                out.append(rb.ravel() - self.K[k] @ vel[k])
                out.append(-self.K[k].T @ l.ravel() + vel[k])
        pos = 0
        for k in self.fiber_rows:
            N = self.sizes[k]
            o = 4 * self.offs[k]
            vk = v_fib[pos:pos + N].T.ravel()
            pos += N
            out.append(self.A_sq[k] @ XT[o:o + 4 * N] - self.Q[k] @ vk)
        return np.concatenate(out) if out else np.zeros(0)
