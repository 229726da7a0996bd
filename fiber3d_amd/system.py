"""GPU-resident implicit hydrodynamic step of free fibers in a confining shell.

Mirrors tstep.time_step_hydro (python/tstep/tstep.py:628-1216) for fibers with
force/torque-free ends inside a spherical periphery:
  * per-step assembly on the GPU (fiber.form_linear_operator / compute_RHS /
    apply_BC_rectangular / force_operator, python/fiber/fiber.py:322-961,
    :1399-1431): the fused f3d_fiber_assemble kernel, one launch per node
    count (assembly='hip', default), or the same formulas as batched torch
    operations (assembly='torch', kept as a second implementation to test
    against);
  * the matvec A_body_fiber_hyd (tstep.py:1003-1178) from libf3d kernels:
    batched block GEMVs for the force operator, the tension projection and
    A_sq XT, the Oseen / stresslet pair sums with same-fiber exclusion
    (replacing the add-then-subtract of the dense self blocks,
    tstep_utils.py:703-729, :812-851) and the matrix-free periphery operator;
  * the block preconditioner P_inv (tstep.py:950-990): batched fiber LU solves
    and a dense shell LU;
  * right-preconditioned FP64 GMRES (gmres.py in this package).
  * free rigid bodies (no fiber attached): surface double layer with
    singularity subtraction, K = [J, rot], the body rows of the matvec
    (tstep.py:1044-1143), external force/torque as a Stokeslet + rotlet at the
    centre (tstep.py:818-846) and a dense per-body preconditioner
    (tstep_utils.py:1341-1387).
Unknown vector layout is the reference's: [shell density (3 Nper, AoS) |
per body density (3 Nb, AoS), U, Omega | per fiber X(N), Y(N), Z(N), T(N)].
"""
import warnings

import numpy as np
import torch
from scipy.linalg import LinAlgWarning

from . import _lib as L
from . import fiber_matrices as FMX
from .dense import BatchedLU, BlockGemv, BlockSet, DenseGemv
from .kernels import _stream, device, pairsum, periphery_apply


def _scatter_blocks(buf, offs_h, blocks):
    """Write equal-length rows blocks[b] at buf[offs_h[b]:...]; a single strided
    copy when the offsets are evenly spaced, else one copy per block."""
    B, n = blocks.shape
    offs_h = np.asarray(offs_h, np.int64)
    if B == 1 or np.all(np.diff(offs_h) == n):
        buf[offs_h[0]:offs_h[0] + B * n].view(B, n).copy_(blocks)
    else:
        for b in range(B):
            buf[offs_h[b]:offs_h[b] + n].copy_(blocks[b])


def dense_inverse(A, panel=64):
    """Inverse of a square matrix from one LU factorisation, solved against
    identity panels of `panel` columns: the triangular-solve workspace hipBLAS
    allocates grows with the right-hand-side width and fails for thousands of
    columns."""
    lu, piv = torch.linalg.lu_factor(A)
    m = A.shape[-1]
    out = torch.empty_like(A)
    for c0 in range(0, m, panel):
        w = min(panel, m - c0)
        e = torch.zeros((m, w), dtype=A.dtype, device=A.device)
        e[c0 + torch.arange(w, device=A.device), torch.arange(w, device=A.device)] = 1.0
        out[:, c0:c0 + w] = torch.linalg.lu_solve(lu, piv, e)
    return out


def _c0c1(epsilon, eta):
    return -np.log(np.e * epsilon**2) / (8.0 * np.pi * eta), 2.0 / (8.0 * np.pi * eta)


class Shell:
    """Static periphery: nodes, inward normals, weights, subtraction vectors and
    the shell block of the preconditioner.

    pc='dense' builds the reference's dense operator (tstep.py:328-345) and
    inverts it once; the reference LU-factors it (tstep.py:346-351) and
    applies lu_solve, the same linear map.  The inverse is applied in FP64 by
    the libf3d dense GEMV: bandwidth-bound instead of a latency-bound pair of
    triangular sweeps, and free of BLAS/rocSOLVER calls, whose workspaces a
    HIP graph cannot keep.  pc_f32=True (the default) keeps an FP32 copy of
    the inverse as well, applied with FP64 arithmetic when the caller asks
    for the FP32 preconditioner (the fast mode, beside its FP32 fiber
    factors): half the HBM traffic per apply.  pc='block' inverts only each
    node's 3x3 diagonal block, -E_i/w_i - I/w_i + n_i n_i^T
    (0.1 ms per apply at 32768 nodes; about twice the GMRES iterations of the
    dense form).  The dense inverse grows as (3N)^2: 4.6 GB and 7 s of set-up
    at 8000 nodes, 77 GB and minutes of set-up at 32768, so pc='auto' picks
    'dense' up to 8192 nodes."""

    DENSE_MAX = 8192

    def __init__(self, x, normals, weights, eta=1.0, reg=5e-3, eps=1e-5, pc="auto",
                 pc_f32=True):
        dev = device()
        self.n = x.shape[0]
        self.pos = torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(dev)
        self.nrm = torch.from_numpy(np.ascontiguousarray(normals, np.float64)).to(dev)
        self.w = torch.from_numpy(np.ascontiguousarray(weights, np.float64)).to(dev)
        self.eta, self.reg, self.eps = eta, reg, eps
        # e_c = stresslet self-sum of w * e_c (periphery.py:75-98)
        self.e = []
        for c in range(3):
            dens = torch.zeros_like(self.pos)
            dens[:, c] = self.w
            self.e.append(pairsum("stresslet", self.pos, dens, self.pos, src_nrm=self.nrm,
                                  eta=eta, reg=reg, eps=eps, exclude_self=True))
        if pc == "auto":
            pc = "dense" if self.n <= self.DENSE_MAX else "block"
        self.pc = pc
        self.minv = self.minv32 = None
        if pc == "dense":
            # FP64 inverse applied with the libf3d dense GEMV (no BLAS call,
            # so the preconditioner can be recorded in a HIP graph)
            self.minv = DenseGemv(dense_inverse(self.dense_matrix()))
            if pc_f32:
                self.minv32 = DenseGemv(self.minv.M.float())
        elif pc == "block":
            E = torch.stack(self.e, dim=2)  # E[i, a, c] = e_c[i, a]
            winv = (1.0 / self.w)[:, None, None]
            I3 = torch.eye(3, dtype=torch.float64, device=dev)[None]
            D = -E * winv - I3 * winv + self.nrm[:, :, None] * self.nrm[:, None, :]
            self.Dinv = torch.linalg.inv(D)
        else:
            raise ValueError("pc is 'auto', 'dense' or 'block'")

    def dense_matrix(self):
        """M of tstep.py:328-345 on the GPU (3N x 3N, AoS ordering)."""
        N = self.n
        M = torch.empty((3 * N, 3 * N), dtype=torch.float64, device=self.pos.device)
        rc = L.load().f3d_stresslet_matrix(N, L.ptr(self.pos), L.ptr(self.nrm), 3, 1,
                                           float(self.eta), float(self.reg), float(self.eps),
                                           L.ptr(M), 3 * N, _stream())
        L.check(rc, "f3d_stresslet_matrix")
        Mv = M.view(N, 3, N, 3)
        idx = torch.arange(N, device=M.device)
        winv = 1.0 / self.w
        for c in range(3):
            # - e_c(i) / w_i in column (i, c); - 1/w_i on the diagonal
            Mv[idx, :, idx, c] -= self.e[c] * winv[:, None]
            Mv[idx, c, idx, c] -= winv
        nflat = self.nrm.reshape(-1)
        M += torch.outer(nflat, nflat)
        return M

    def apply(self, q, add=None, out=None, precision="f64"):
        return periphery_apply(self.pos, self.nrm, self.w, self.e[0], self.e[1], self.e[2], q,
                               add=add, eta=self.eta, reg=self.reg, eps=self.eps,
                               precision=precision, out=out)

    def precond(self, r, out, f32=False):
        """Shell block of P^-1: the dense inverse (an HBM-bound GEMV), from its
        FP32 copy when f32 is set and the copy exists, else in FP64."""
        if self.pc == "block":
            # per node: out_i = D_i^-1 r_i, written as elementwise products
            D, rv = self.Dinv, r.view(self.n, 3)
            out.view(self.n, 3).copy_(D[:, :, 0] * rv[:, 0:1] + D[:, :, 1] * rv[:, 1:2]
                                      + D[:, :, 2] * rv[:, 2:3])
        else:
            src = r if (r.is_contiguous() and r.data_ptr() != out.data_ptr()) else r.clone()
            (self.minv32 if f32 and self.minv32 is not None else self.minv)(src, out)
        return out


class Bodies:
    """Free rigid bodies on the GPU: surface nodes, outward normals, weights,
    singularity-subtraction vectors, K = [J, rot] (body/body.py:162-193,
    :325-346) and the dense per-body preconditioner block
    A = [M - I_sub, -K; -K^T, I] (tstep_utils.py:1341-1387).

    The preconditioner follows the reference's precompute option
    (tstep.py:225-252, :966-980): A is inverted once, in each body's initial
    pose, and applied in the body frame.  A rigid rotation R of a body maps
    its block to Q A0 Q^T with Q = diag(R per node, R, R) (the stresslet is a
    rank-3 tensor, K picks up R on both sides), so A^-1 = Q A0^-1 Q^T: every
    3-vector of the body's slice is rotated by R^T, multiplied by A0^-1
    (one batched FP64 GEMV) and rotated back.  All bodies must carry the same
    node count (batched apply)."""

    def __init__(self, bodies, eta, reg, eps, body_force=None, orientations=None):
        dev = device()
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        sizes = [np.asarray(b[1]).shape[0] for b in bodies]
        if len(set(sizes)) != 1:
            raise ValueError("bodies must have equal node counts")
        self.count = len(bodies)
        self.nb = sizes[0]
        self.n = self.count * self.nb
        self.eta, self.reg, self.eps = eta, reg, eps
        self.centres = t(np.array([b[0] for b in bodies]))
        self.pos = t(np.concatenate([b[1] for b in bodies]))
        self.nrm = t(np.concatenate([b[2] for b in bodies]))
        self.w = t(np.concatenate([b[3] for b in bodies]))
        self.bid = torch.arange(self.count, device=dev).repeat_interleave(self.nb)
        f = np.zeros((self.count, 6)) if body_force is None else np.asarray(body_force, float)
        self.force = t(f)
        self.vb = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)
        # rotation of every body from the pose its inverse block was built in
        self.rot = torch.eye(3, dtype=torch.float64, device=dev).repeat(self.count, 1, 1)
        # orientation quaternion (s, p) of the reference's bodies, advanced
        # like body.orientation (tstep.py:1404): q <- q(Omega dt) q
        q0 = np.tile([1.0, 0.0, 0.0, 0.0], (self.count, 1)) if orientations is None \
            else np.asarray(orientations, float).reshape(self.count, 4)
        self.quat = t(q0 / np.linalg.norm(q0, axis=1, keepdims=True))
        self.update_geometry()
        self.ainv = DenseGemv(torch.stack([dense_inverse(A) for A in self.pc_blocks()]))
        m = 3 * self.nb + 6
        self._xb = torch.empty(self.count * m, dtype=torch.float64, device=dev)
        self._yb = torch.empty(self.count * m, dtype=torch.float64, device=dev)

    def update_geometry(self):
        """Centre-relative nodes and subtraction vectors for the current pose."""
        nb = self.nb
        r = self.pos - self.centres[self.bid]
        if hasattr(self, "r"):
            self.r.copy_(r)
        else:
            self.r = r
        e_new = []
        for c in range(3):
            dens = torch.zeros_like(self.pos)
            dens[:, c] = self.w
            e = torch.empty_like(self.pos)
            for k in range(self.count):
                s = slice(k * nb, (k + 1) * nb)
                pairsum("stresslet", self.pos[s], dens[s], self.pos[s], src_nrm=self.nrm[s],
                        eta=self.eta, reg=self.reg, eps=self.eps, out=e[s], exclude_self=True)
            e_new.append(e)
        if hasattr(self, "e"):
            for c in range(3):
                self.e[c].copy_(e_new[c])
        else:
            self.e = e_new

    def pc_blocks(self):
        """The dense blocks [M - I_sub, -K; -K^T, I] of the current pose."""
        nb = self.nb
        m = 3 * nb + 6
        A = torch.zeros((self.count, m, m), dtype=torch.float64, device=self.pos.device)
        for k in range(self.count):
            s = slice(k * nb, (k + 1) * nb)
            Mk = A[k, :3 * nb, :3 * nb]
            rc = L.load().f3d_stresslet_matrix(nb, L.ptr(self.pos[s]), L.ptr(self.nrm[s]), 3, 1,
                                               float(self.eta), float(self.reg), float(self.eps),
                                               L.ptr(A[k]), m, _stream())
            L.check(rc, "f3d_stresslet_matrix")
            Mv = Mk.view(nb, 3, nb, 3)
            idx = torch.arange(nb, device=A.device)
            winv = 1.0 / self.w[s]
            for c in range(3):
                Mv[idx, :, idx, c] -= self.e[c][s] * winv[:, None]
            Kb = self.k_matrix(self.r[s])
            A[k, :3 * nb, 3 * nb:] = -Kb
            A[k, 3 * nb:, :3 * nb] = -Kb.t()
            A[k, 3 * nb:, 3 * nb:] = torch.eye(6, dtype=torch.float64, device=A.device)
        return A

    @staticmethod
    def k_matrix(r):
        n = r.shape[0]
        K = torch.zeros((n, 3, 6), dtype=torch.float64, device=r.device)
        K[:, 0, 0] = K[:, 1, 1] = K[:, 2, 2] = 1.0
        x, y, z = r[:, 0], r[:, 1], r[:, 2]
        K[:, 0, 4], K[:, 0, 5] = z, -y
        K[:, 1, 3], K[:, 1, 5] = -z, x
        K[:, 2, 3], K[:, 2, 4] = y, -x
        return K.reshape(3 * n, 6)

    def block(self, vec, start):
        """(count, 3 Nb + 6) view of the body unknowns: per body the density
        (AoS) followed by U and Omega, as in the reference's layout
        (tstep.py:1017-1025)."""
        m = 3 * self.nb + 6
        return vec[start:start + self.count * m].view(self.count, m)

    def density_view(self, x, start):
        """Contiguous (n, 3) copy of all body densities."""
        return self.block(x, start)[:, :3 * self.nb].reshape(self.n, 3)

    def _vel(self, x, start):
        return self.block(x, start)[:, 3 * self.nb:]

    def explicit(self, trg, out, accumulate=True):
        """Stokeslet + rotlet of the external force/torque at each centre."""
        kw = dict(eta=self.eta, reg=self.reg, eps=self.eps)
        pairsum("oseen", self.centres, self.force[:, :3].contiguous(), trg, out=out,
                accumulate=accumulate, **kw)
        pairsum("rotlet", self.centres, self.force[:, 3:].contiguous(), trg, out=out,
                accumulate=True, **kw)
        return out

    def scatter_density(self, vec, v, start):
        """Write per-node 3-vectors v (n, 3) into the density slots of vec."""
        self.block(vec, start)[:, :3 * self.nb] = v.reshape(self.count, 3 * self.nb)

    def rows(self, x, out, start):
        """Body density rows vb - sum_c (lam_c / w) e_c - K U and velocity rows
        -K^T lam + U (tstep.py:1121-1143)."""
        lam = self.density_view(x, start)
        vel = self._vel(x, start)
        a = lam / self.w[:, None]
        sub = a[:, 0:1] * self.e[0] + a[:, 1:2] * self.e[1] + a[:, 2:3] * self.e[2]
        U = vel[self.bid, :3]
        W = vel[self.bid, 3:]
        KU = U + torch.cross(W, self.r, dim=1)
        ob = self.block(out, start)
        ob[:, :3 * self.nb] = (self.vb - sub - KU).reshape(self.count, 3 * self.nb)
        F = torch.zeros((self.count, 3), dtype=torch.float64, device=x.device).index_add_(0, self.bid, lam)
        T = torch.zeros_like(F).index_add_(0, self.bid, torch.cross(self.r, lam, dim=1))
        ob[:, 3 * self.nb:3 * self.nb + 3] = vel[:, :3] - F
        ob[:, 3 * self.nb + 3:] = vel[:, 3:] - T

    def precond(self, x, out, start):
        """Rotate every 3-vector of a body's slice into the body frame
        ((R^T v)_c = sum_a R_ac v_a), apply A0^-1 (batched dense GEMV),
        rotate back ((R y)_c = sum_a R_ca y_a).  Elementwise products and the
        libf3d GEMV only, no BLAS, so the step can be recorded in a graph."""
        k = self.nb + 2
        R = self.rot[:, None]                              # (count, 1, 3, 3)
        v = self.block(x, start).view(self.count, k, 3)
        xb = self._xb.view(self.count, k, 3)
        xb.copy_(v[..., 0:1] * R[..., 0, :] + v[..., 1:2] * R[..., 1, :] + v[..., 2:3] * R[..., 2, :])
        self.ainv(self._xb, self._yb)
        y = self._yb.view(self.count, k, 3)
        self.block(out, start).view(self.count, k, 3).copy_(
            y[..., 0:1] * R[..., :, 0] + y[..., 1:2] * R[..., :, 1] + y[..., 2:3] * R[..., :, 2])

    def state(self):
        """Copies of the pose (centres, nodes, normals, rotation, quaternion)."""
        return tuple(t.clone() for t in (self.centres, self.pos, self.nrm, self.rot, self.quat))

    def set_state(self, st):
        """Restore a pose taken by state(), in place."""
        for t, v in zip((self.centres, self.pos, self.nrm, self.rot, self.quat), st):
            t.copy_(v)
        self.update_geometry()

    def move(self, vel, dt):
        """Rigid update: centre += U dt, nodes rotated by Omega dt (Rodrigues)."""
        U, W = vel[:, :3], vel[:, 3:]
        th = torch.linalg.norm(W, dim=1) * dt
        ax = W / torch.linalg.norm(W, dim=1, keepdim=True).clamp_min(1e-300)
        c, s = torch.cos(th)[:, None], torch.sin(th)[:, None]
        # R = c I + s [k]_x + (1 - c) k k^T per body, composed onto self.rot
        kx = torch.zeros((self.count, 3, 3), dtype=torch.float64, device=W.device)
        kx[:, 0, 1], kx[:, 0, 2], kx[:, 1, 2] = -ax[:, 2], ax[:, 1], -ax[:, 0]
        kx = kx - kx.transpose(1, 2)
        Rinc = c[:, :, None] * torch.eye(3, dtype=torch.float64, device=W.device) \
            + s[:, :, None] * kx + (1 - c)[:, :, None] * ax[:, :, None] * ax[:, None, :]
        self.rot.copy_(torch.bmm(Rinc, self.rot))
        # q <- (cos(th/2), sin(th/2) axis) q, Hamilton product (quaternion.py:16-40)
        qs, qp = torch.cos(th / 2)[:, None], torch.sin(th / 2)[:, None] * ax
        s0, p0 = self.quat[:, :1], self.quat[:, 1:]
        self.quat.copy_(torch.cat([qs * s0 - (qp * p0).sum(1, keepdim=True),
                                   qs * p0 + s0 * qp + torch.cross(qp, p0, dim=1)], dim=1))
        def rot(v):
            k = ax[self.bid]
            cc, ss = c[self.bid], s[self.bid]
            return v * cc + torch.cross(k, v, dim=1) * ss + k * (k * v).sum(1, keepdim=True) * (1 - cc)
        # in place: recorded graphs keep the addresses of these tensors
        self.centres.add_(U * dt)
        self.pos.copy_(self.centres[self.bid] + rot(self.r))
        self.nrm.copy_(rot(self.nrm))
        self.update_geometry()


class FiberSystem:
    def __init__(self, fibers, shell=None, eta=1.0, dt=1e-3, E=0.01, epsilon=1e-3,
                 penalty=500.0, nfd=4, end_force=1.0, discretisation="reference",
                 lengths=None, pair_precision="f64", lu_precision="f64", lu_refine=6,
                 lu_tol=1e-13, reg=5e-3, eps=1e-5, bodies=None, body_force=None,
                 assembly="hip", attached=None, body_orientations=None, hinged=None,
                 cortex=False):
        self.dev = device()
        self.eta, self.dt, self.E, self.penalty = eta, dt, E, penalty
        self.c0, self.c1 = _c0c1(epsilon, eta)
        self.nfd, self.kind = nfd, discretisation
        self.reg, self.eps = reg, eps
        self.end_force = end_force
        self.pair_precision = pair_precision
        if assembly not in ("torch", "hip"):
            raise ValueError("assembly is 'torch' or 'hip'")
        self.assembly = assembly
        self.lu_precision, self.lu_refine, self.lu_tol = lu_precision, lu_refine, lu_tol
        sizes = np.array([np.asarray(x).shape[0] for x in fibers], dtype=np.int64)
        if sizes.min() < 5:
            raise ValueError("fibers need at least 5 nodes")
        self.sizes = sizes
        self.nfib = len(sizes)
        self.offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.npts = int(self.offs[-1])
        self.shell = shell
        self.nper = 0 if shell is None else shell.n
        self.bodies = Bodies(bodies, eta, reg, eps, body_force, body_orientations) \
            if bodies else None
        self.nbody = 0 if self.bodies is None else self.bodies.count
        self.nbn = 0 if self.bodies is None else self.bodies.n
        self.fib0 = 3 * self.nper + 3 * self.nbn + 6 * self.nbody
        self.size = self.fib0 + 4 * self.npts
        self.lengths_h = np.full(self.nfib, 1.0) if lengths is None else np.asarray(lengths, float)
        self.lengths = torch.from_numpy(self.lengths_h).to(self.dev)
        # node positions (npts, 3), AoS
        self.X = torch.from_numpy(np.ascontiguousarray(np.concatenate(fibers), np.float64)).to(self.dev)
        # buckets of equal node count
        self.buckets = {}
        for N in sorted(set(sizes.tolist())):
            ids = np.flatnonzero(sizes == N)
            fm = FMX.get(int(N), nfd, discretisation)
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(self.dev)
            self.buckets[int(N)] = {
                "ids": ids, "ids_t": torch.from_numpy(ids).to(self.dev),
                "D0": [t(D) for D in fm.D0], "P": t(fm.P_cheb), "w0": t(fm.weights0),
                "alpha": t(fm.alpha),
                # node index of each bucket fiber's first node
                "start": torch.from_numpy(self.offs[ids]).to(self.dev),
            }
        # same-fiber exclusion ranges for fiber targets
        lo = np.repeat(self.offs[:-1], sizes).astype(np.int32)
        hi = np.repeat(self.offs[1:], sizes).astype(np.int32)
        self.excl = (torch.from_numpy(lo).to(self.dev), torch.from_numpy(hi).to(self.dev))
        # plus-end node of every fiber, and for fiber targets the exclusion of
        # their own fiber's end node in the list of end nodes
        self.end_idx = torch.from_numpy(self.offs[1:] - 1).to(self.dev)
        fid = np.repeat(np.arange(self.nfib), sizes).astype(np.int32)
        self.excl_end = (torch.from_numpy(fid).to(self.dev), torch.from_numpy(fid + 1).to(self.dev))
        # block layouts
        self.bs = BlockSet(4 * sizes)  # A_sq blocks (4N x 4N)
        fo = self.offs[:-1]
        # force operator: 3 row blocks per fiber (N x 4N) -> SoA f (3, npts)
        mF = np.repeat(sizes, 3)
        kF = 4 * mF
        xF = np.repeat(4 * fo, 3)
        yF = (np.arange(3)[None, :] * self.npts + fo[:, None]).ravel()
        self.gF = BlockGemv(mF, kF, xF, yF)
        # tension projection: per component c, (4N-14) x N blocks
        self.gQ = [BlockGemv(4 * sizes - 14, sizes, c * self.npts + fo, 4 * fo) for c in range(3)]
        self.gA = BlockGemv(4 * sizes, 4 * sizes, 4 * fo, 4 * fo)
        # launch tables of the fused assembly kernel, per node count
        ti = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int64)).to(self.dev)
        f_off_all = self.gF.m_off_h.reshape(self.nfib, 3)[:, 0]
        for N, bk in self.buckets.items():
            ids = bk["ids"]
            P = bk["P"]
            bk["asm"] = {
                "node0": ti(self.offs[ids]), "a_off": ti(self.bs.a_off_h[ids]),
                "f_off": ti(f_off_all[ids]),
                "q_off": ti(np.stack([self.gQ[c].m_off_h[ids] for c in range(3)], axis=1)),
                "D0": torch.stack(bk["D0"]).contiguous(),
                "PX": P[:N - 4, :N].contiguous(), "PT": P[3 * (N - 4):, 3 * N:].contiguous(),
                "length": self.lengths[bk["ids_t"]].contiguous(),
                "rhs_off": ti(self.fib0 + 4 * self.offs[ids]),
            }
        self._setup_attached(attached)
        self._setup_cortex(cortex)
        self._setup_hinged(hinged)
        # work buffers
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=self.dev)
        self.f_soa = z(3 * self.npts)
        self.v_soa = z(3 * self.npts)
        self.v_shell = torch.zeros((max(self.nper, 1), 3), dtype=torch.float64, device=self.dev)
        self.prepare()

    # ------------------------------------------------------------------
    def _setup_attached(self, attached):
        """Fibers whose minus end is clamped to a body (tstep.py:878-881):
        attached = [(fiber k, body j, link (3,))], the link being the
        nucleation site relative to the body centre in the body's pose at
        construction.  Grouped by node count for the batched row patches."""
        self.att = {}
        if not attached:
            return
        if self.bodies is None:
            raise ValueError("attached fibers need bodies")
        ak = np.array([int(a[0]) for a in attached])
        aj = np.array([int(a[1]) for a in attached])
        lref = np.array([np.asarray(a[2], float) for a in attached]).reshape(-1, 3)
        if len(set(ak.tolist())) != ak.size or ak.min() < 0 or ak.max() >= self.nfib:
            raise ValueError("attached: fiber indices must be distinct and in range")
        if aj.min() < 0 or aj.max() >= self.nbody:
            raise ValueError("attached: body index out of range")
        if np.any(np.linalg.norm(lref, axis=1) == 0.0):
            raise ValueError("attached: a link must not sit at the body centre")
        ti = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int64)).to(self.dev)
        for N, bk in self.buckets.items():
            sel = np.flatnonzero(np.isin(ak, bk["ids"]))
            if sel.size == 0:
                continue
            fib = ak[sel]
            self.att[N] = {
                "fib": fib, "pib": ti(np.searchsorted(bk["ids"], fib)), "body": ti(aj[sel]),
                "lref": torch.from_numpy(lref[sel]).to(self.dev),
                "node0": ti(self.offs[fib]), "xo": ti(4 * self.offs[fib]),
                "a_off": ti(self.bs.a_off_h[fib]),
                "row0": ti(self.fib0 + 4 * self.offs[fib] + 4 * N - 14),
            }

    def _setup_cortex(self, cortex):
        """attached_to_cortex (tstep.py:881-883): every fiber not attached to
        a body has its minus end clamped in place, the same rows as a fiber
        clamped to a body at rest, with no link."""
        self.cor = {}
        if not cortex:
            return
        taken = set()
        for g in self.att.values():
            taken |= set(g["fib"].tolist())
        ti = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int64)).to(self.dev)
        for N, bk in self.buckets.items():
            fib = np.array([k for k in bk["ids"] if k not in taken], dtype=np.int64)
            if fib.size == 0:
                continue
            self.cor[N] = {
                "fib": fib, "pib": ti(np.searchsorted(bk["ids"], fib)),
                "node0": ti(self.offs[fib]), "a_off": ti(self.bs.a_off_h[fib]),
                "row0": ti(self.fib0 + 4 * self.offs[fib] + 4 * N - 14),
            }

    def _setup_hinged(self, hinged):
        """Fibers whose plus end is held in place, free to rotate, as a fiber
        that reached the periphery (tstep.py:886-888)."""
        self.hng = {}
        if not hinged:
            return
        hk = np.unique(np.asarray(hinged, dtype=np.int64))
        if hk.min() < 0 or hk.max() >= self.nfib:
            raise ValueError("hinged: fiber index out of range")
        ti = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int64)).to(self.dev)
        for N, bk in self.buckets.items():
            fib = hk[np.isin(hk, bk["ids"])]
            if fib.size == 0:
                continue
            self.hng[N] = {
                "fib": fib, "pib": ti(np.searchsorted(bk["ids"], fib)),
                "nodeL": ti(self.offs[fib + 1] - 1), "a_off": ti(self.bs.a_off_h[fib]),
                "row0": ti(self.fib0 + 4 * self.offs[fib] + 4 * N - 7),
            }

    def _apply_hinged(self):
        """Replace the plus-end force rows of every hinged fiber by the
        'velocity' rows at s = L (fiber.py:940-960): x(L)/dt and the tension
        row 6 E c0 x_ss(L).D3 X + 2 c0 D1 T, right-hand side
        -(v(L) + 2 c0 f(L)).x_s(L); the torque rows stay."""
        dt, E, c0 = self.dt, self.E, self.c0
        for N, g in self.hng.items():
            node, X, xs, xss, D1, P, B = self._ctx[N]
            bk = self.buckets[N]
            sc = 2.0 / self.lengths[bk["ids_t"][g["pib"]]]
            d1 = bk["D0"][0][-1][None, :] * sc[:, None]
            d3 = bk["D0"][2][-1][None, :] * sc[:, None] ** 3
            xsL, xssL = xs[g["pib"], -1], xss[g["pib"], -1]
            na, n = len(g["fib"]), 4 * N
            R = torch.zeros((na, 4, n), dtype=torch.float64, device=self.dev)
            for a in range(3):
                R[:, a, a * N + N - 1] = 1.0 / dt
                R[:, 3, a * N:(a + 1) * N] = 6.0 * E * c0 * xssL[:, a, None] * d3
            R[:, 3, 3 * N:] = 2.0 * c0 * d1
            rows = torch.arange(n - 7, n - 3, device=self.dev)
            cols = torch.arange(n, device=self.dev)
            idx = g["a_off"][:, None, None] + cols[None, None, :] * n + rows[None, :, None]
            self.A_buf[idx.reshape(-1)] = R.reshape(-1)
            r0 = g["row0"]
            XL = self.X[g["nodeL"]]
            vL, fL = self.v_explicit[g["nodeL"]], self.force[g["nodeL"]]
            for a in range(3):
                self.rhs[r0 + a] = XL[:, a] / dt
            self.rhs[r0 + 3] = -((vL + 2.0 * c0 * fL) * xsL).sum(1)
            # refilled in place: a recorded Arnoldi graph keeps this address
            if "xsL" not in g:
                g["xsL"] = xsL.clone()
            else:
                g["xsL"].copy_(xsL)

    def _links(self, g):
        """Current link vectors (rotated with their bodies) and directions."""
        link = torch.einsum("nij,nj->ni", self.bodies.rot[g["body"]], g["lref"])
        return link, link / torch.linalg.norm(link, dim=1, keepdim=True)

    def _apply_clamped(self):
        """Replace the minus-end boundary rows of every fiber clamped to a
        body or to the cortex by the 'velocity' and 'angular_velocity' rows (fiber.py:855-877, :887-898):
        x(0)/dt, x_s(0)/dt, and the tension row 6 E c0 x_ss(0).D3 X +
        2 c0 D1 T with right-hand side -(v(0) + 2 c0 f(0)).x_s(0)."""
        dt, E, c0 = self.dt, self.E, self.c0
        for N, g in list(self.att.items()) + list(self.cor.items()):
            node, X, xs, xss, D1, P, B = self._ctx[N]
            bk = self.buckets[N]
            sc = 2.0 / self.lengths[bk["ids_t"][g["pib"]]]
            d1 = bk["D0"][0][0][None, :] * sc[:, None]
            d2 = bk["D0"][1][0][None, :] * sc[:, None] ** 2
            d3 = bk["D0"][2][0][None, :] * sc[:, None] ** 3
            xs0, xss0 = xs[g["pib"], 0], xss[g["pib"], 0]
            na, n = len(g["fib"]), 4 * N
            R = torch.zeros((na, 7, n), dtype=torch.float64, device=self.dev)
            for a in range(3):
                R[:, a, a * N] = 1.0 / dt
                R[:, 3, a * N:(a + 1) * N] = 6.0 * E * c0 * xss0[:, a, None] * d3
                R[:, 4 + a, a * N:(a + 1) * N] = d1 / dt
            R[:, 3, 3 * N:] = 2.0 * c0 * d1
            # column-major blocks: entry (row, col) at a_off + col * n + row
            rows = torch.arange(n - 14, n - 7, device=self.dev)
            cols = torch.arange(n, device=self.dev)
            idx = g["a_off"][:, None, None] + cols[None, None, :] * n + rows[None, :, None]
            self.A_buf[idx.reshape(-1)] = R.reshape(-1)
            r0 = g["row0"]
            X0 = self.X[g["node0"]]
            v0, f0 = self.v_explicit[g["node0"]], self.force[g["node0"]]
            for a in range(3):
                self.rhs[r0 + a] = X0[:, a] / dt
                self.rhs[r0 + 4 + a] = xs0[:, a] / dt
            self.rhs[r0 + 3] = -((v0 + 2.0 * c0 * f0) * xs0).sum(1)
            # refilled in place: a recorded Arnoldi graph keeps these addresses
            if "xs0" not in g:
                g["xs0"], g["d2"], g["d3"] = xs0.clone(), d2, d3
            else:
                g["xs0"].copy_(xs0)

    def _link_forces(self, XT):
        """Force and torque of the attached fibers on their bodies (the body
        rows of tstep_utils.build_link_matrix, tstep_utils.py:995-1036):
        F = -E X_sss(0) + T(0) x_s(0), L = link x F + E x_s(0) x X_ss(0),
        summed per body -> (nbody, 6)."""
        ft = torch.zeros((self.nbody, 6), dtype=torch.float64, device=self.dev)
        for N, g in self.att.items():
            blk = XT[g["xo"][:, None] + torch.arange(4 * N, device=self.dev)[None, :]]
            Xc = blk[:, :3 * N].view(-1, 3, N)
            T0 = blk[:, 3 * N]
            F = -self.E * (Xc * g["d3"][:, None, :]).sum(-1) + T0[:, None] * g["xs0"]
            link, _ = self._links(g)
            tau = torch.cross(link, F, dim=1) \
                + self.E * torch.cross(g["xs0"], (Xc * g["d2"][:, None, :]).sum(-1), dim=1)
            ft.index_add_(0, g["body"], torch.cat([F, tau], dim=1))
        return ft

    def _link_rows(self, x, out, vsoa):
        """The fiber rows of the link matrix (tstep_utils.py:1039-1088):
        -(U + Omega x link) in the position rows, -(U + Omega x link).x_s(0)
        plus the flow term v(0).x_s(0) (tstep.py:1162-1166) in the tension
        row, -Omega x link_dir in the orientation rows."""
        vel = self.bodies._vel(x, 3 * self.nper)
        for N, g in self.att.items():
            U, W = vel[g["body"], :3], vel[g["body"], 3:]
            link, ldir = self._links(g)
            vU = U + torch.cross(W, link, dim=1)
            xs0 = g["xs0"]
            r0 = g["row0"]
            add = torch.cat([-vU, (-(vU * xs0) + vsoa[g["node0"]] * xs0).sum(1, keepdim=True),
                             -torch.cross(W, ldir, dim=1)], dim=1)
            idx = r0[:, None] + torch.arange(7, device=self.dev)[None, :]
            out.index_add_(0, idx.reshape(-1), add.reshape(-1))

    def fix_links(self):
        """Translate every attached fiber so that its minus end sits on its
        body's (rotated) nucleation site (tstep.py:462-470)."""
        for N, g in self.att.items():
            link, _ = self._links(g)
            shift = self.X[g["node0"]] - (self.bodies.centres[g["body"]] + link)
            nodes = g["node0"][:, None] + torch.arange(N, device=self.dev)[None, :]
            self.X.index_add_(0, nodes.reshape(-1), -shift.repeat_interleave(N, dim=0))

    def _soa(self, buf):
        """(npts, 3) view of an SoA (3, npts) buffer."""
        return buf.view(3, self.npts).t()

    def prepare(self, refactor=True):
        """Per-step assembly: derivatives, A_sq, RHS, force operator, tension
        projection, explicit flows, LU factors (tstep.py:689-941).

        refactor=False keeps the fiber LU factors of an earlier assembly as the
        preconditioner (a lagged preconditioner: GMRES still solves with the
        current operator, and the current A_sq is what any refinement uses)."""
        dev = self.dev
        E, dt, c0, c1, p = self.E, self.dt, self.c0, self.c1, self.penalty
        # Every buffer the matvec and the preconditioner read is allocated once
        # and re-filled in place, so a recorded HIP graph of the Arnoldi
        # operator stays valid from one step to the next.
        if not hasattr(self, "A_buf"):
            self.A_buf = torch.empty(self.bs.nelem, dtype=torch.float64, device=dev)
            self.F_buf = torch.empty(self.gF.nelem, dtype=torch.float64, device=dev)
            self.Q_buf = [torch.empty(g.nelem, dtype=torch.float64, device=dev) for g in self.gQ]
            self.weights = torch.empty(self.npts, dtype=torch.float64, device=dev)
            self.xs = torch.empty((self.npts, 3), dtype=torch.float64, device=dev)
            self.xss = torch.empty((self.npts, 3), dtype=torch.float64, device=dev)
        A_buf, F_buf, Q_buf = self.A_buf, self.F_buf, self.Q_buf
        self._rhs_parts = {}
        self._ctx_data = {}
        if self.assembly == "hip":
            self._assemble_hip(A_buf, F_buf, Q_buf)
            self._ctx_data = None  # per-bucket views, formed on first use
            buckets = {}
        else:
            buckets = self.buckets
        for N, bk in buckets.items():
            ids_t = bk["ids_t"]
            B = ids_t.numel()
            node = bk["start"][:, None] + torch.arange(N, device=dev)[None, :]  # (B, N)
            X = self.X[node]  # (B, N, 3)
            sc = (2.0 / self.lengths[ids_t])[:, None, None]
            Ds = [bk["D0"][k][None] * sc ** (k + 1) for k in range(4)]  # (B, N, N)
            xs, xss, xsss = (torch.matmul(Ds[k], X) for k in range(3))
            I = torch.eye(N, dtype=torch.float64, device=dev)[None]
            D1, D2, D3, D4 = Ds
            dl = lambda v, D: v[:, :, None] * D
            diag = lambda v: torch.diag_embed(v)
            A = torch.zeros((B, 4 * N, 4 * N), dtype=torch.float64, device=dev)
            for a in range(3):
                for b in range(a, 3):
                    if a == b:
                        blk = I / dt + E * c0 * dl(1.0 + xs[..., a] ** 2, D4) \
                            + E * c1 * dl(1.0 - xs[..., a] ** 2, D4)
                    else:
                        blk = E * (c0 - c1) * dl(xs[..., a] * xs[..., b], D4)
                    A[:, a * N:(a + 1) * N, b * N:(b + 1) * N] = blk
                    A[:, b * N:(b + 1) * N, a * N:(a + 1) * N] = blk
                A[:, a * N:(a + 1) * N, 3 * N:] = -c0 * (2.0 * dl(xs[..., a], D1) + diag(xss[..., a])) \
                    - c1 * diag(xss[..., a])
                A[:, 3 * N:, a * N:(a + 1) * N] = -(c1 + 7.0 * c0) * E * dl(xss[..., a], D4) \
                    - 6.0 * c0 * E * dl(xsss[..., a], D3) - p * dl(xs[..., a], D1)
            A[:, 3 * N:, 3 * N:] = -2.0 * c0 * D2 + (c1 + c0) * diag((xss**2).sum(-1))
            # boundary rows (force/torque-free ends, fiber.py:836-959)
            Brow = torch.zeros((B, 14, 4 * N), dtype=torch.float64, device=dev)
            for a in range(3):
                Brow[:, a, a * N:(a + 1) * N] = E * D3[:, 0]
                Brow[:, a, 3 * N] = -xs[:, 0, a]
                Brow[:, 3, a * N:(a + 1) * N] = -E * D2[:, 0] * xss[:, 0, a, None]
                Brow[:, 4 + a, a * N:(a + 1) * N] = D2[:, 0]
                Brow[:, 7 + a, a * N:(a + 1) * N] = -E * D3[:, -1]
                Brow[:, 7 + a, 4 * N - 1] = xs[:, -1, a]
                Brow[:, 10, a * N:(a + 1) * N] = E * D2[:, -1] * xss[:, -1, a, None]
                Brow[:, 11 + a, a * N:(a + 1) * N] = D2[:, -1]
            Brow[:, 3, 3 * N] = -1.0
            Brow[:, 10, 4 * N - 1] = 1.0
            A_sq = torch.cat((torch.matmul(bk["P"], A), Brow), dim=1)  # (B, 4N, 4N)
            # column-major blocks into the global buffer at their fiber offsets
            _scatter_blocks(A_buf, self.bs.a_off_h[bk["ids"]], A_sq.transpose(1, 2).reshape(B, -1))
            # force operator rows c: [-E D4 in block c | diag(xss_c) + diag(xs_c) D1]
            Fb = torch.zeros((B, 3, N, 4 * N), dtype=torch.float64, device=dev)
            for a in range(3):
                Fb[:, a, :, a * N:(a + 1) * N] = -E * D4
                Fb[:, a, :, 3 * N:] = diag(xss[..., a]) + dl(xs[..., a], D1)
            f_off = self.gF.m_off_h.reshape(self.nfib, 3)[bk["ids"]]
            # the three row blocks of a fiber are adjacent: one (B, 3*N*4N) copy
            _scatter_blocks(F_buf, f_off[:, 0], Fb.transpose(2, 3).reshape(B, -1))
            # tension projection Q_c = P_cheb [e_c ; diag(xs_c) D1] (4N-14 x N)
            P = bk["P"]
            for a in range(3):
                Tc = torch.zeros((B, 4 * N, N), dtype=torch.float64, device=dev)
                Tc[:, a * N:(a + 1) * N] = I
                Tc[:, 3 * N:] = dl(xs[..., a], D1)
                Qc = torch.matmul(P, Tc)  # (B, 4N-14, N)
                _scatter_blocks(Q_buf[a], self.gQ[a].m_off_h[bk["ids"]],
                                Qc.transpose(1, 2).reshape(B, -1))
            w = bk["w0"][None, :] * (self.lengths[ids_t][:, None] / 2.0)
            self.weights[node.reshape(-1)] = w.reshape(-1)
            self.xs[node.reshape(-1)] = xs.reshape(-1, 3)
            self._ctx_data[N] = (node, X, xs, xss, D1, P, B)
        self._build_rhs()
        if self.att or self.cor:
            self._apply_clamped()
        if self.hng:
            self._apply_hinged()
        factored = True
        if getattr(self, "lu", None) is None:
            self.lu = BatchedLU(self.bs, self.A_buf, precision=self.lu_precision,
                                refine=self.lu_refine, tol=self.lu_tol)
        elif refactor:
            self.lu.factor()
        else:
            factored = False
        if factored and int(self.lu.info.max()) > 0:
            # scipy.linalg.lu_factor warns on an exactly-zero pivot; one device
            # read per factorisation
            bad = self.lu.singular_blocks()
            warnings.warn(f"{len(bad)} fiber block(s) exactly singular (fiber, first zero "
                          f"pivot): {bad[:8]}", LinAlgWarning, stacklevel=2)

    def _assemble_hip(self, A_buf, F_buf, Q_buf):
        """A_sq, force rows, tension projections, tangents and weights of every
        fiber from f3d_fiber_assemble, one launch per node count."""
        lib = L.load()
        for N, bk in self.buckets.items():
            t = bk["asm"]
            B = len(bk["ids"])
            rc = lib.f3d_fiber_assemble(
                B, N, L.ptr(t["node0"]), L.ptr(t["length"]), L.ptr(t["a_off"]), L.ptr(t["f_off"]),
                L.ptr(t["q_off"]), L.ptr(self.X), L.ptr(t["D0"]), L.ptr(t["PX"]), L.ptr(t["PT"]),
                L.ptr(bk["w0"]), float(self.E), float(self.dt), float(self.c0), float(self.c1),
                float(self.penalty), L.ptr(A_buf), L.ptr(F_buf), L.ptr(Q_buf[0]), L.ptr(Q_buf[1]),
                L.ptr(Q_buf[2]), L.ptr(self.xs), L.ptr(self.xss), L.ptr(self.weights), _stream())
            L.check(rc, "f3d_fiber_assemble")

    @property
    def _ctx(self):
        """Per node count N: (node indices (B, N), X, xs, xss at them (B, N, 3),
        scaled D_1 (B, N, N), projection P, fiber count B) of the current
        assembly; the HIP path forms them only when something asks."""
        if self._ctx_data is None:
            ctx = {}
            for N, bk in self.buckets.items():
                if "node" not in bk:
                    bk["node"] = bk["start"][:, None] + torch.arange(N, device=self.dev)[None, :]
                    sc = (2.0 / self.lengths[bk["ids_t"]])[:, None, None]
                    bk["D1"] = bk["D0"][0][None] * sc
                node = bk["node"]
                ctx[N] = (node, self.X[node], self.xs[node], self.xss[node], bk["D1"], bk["P"],
                          len(bk["ids"]))
            self._ctx_data = ctx
        return self._ctx_data

    def _build_rhs(self):
        """RHS: explicit flows of the plus-end forces and per-fiber RHS
        (tstep.py:794-846, tstep_utils.py:1238-1336, fiber.py:473-582)."""
        dev = self.dev
        # explicit force: plus end along the end tangent.  Only the end nodes
        # carry force, so only they are sources of the explicit flows (the
        # zero-force nodes add exact zeros to the reference's all-node sums).
        force = torch.zeros((self.npts, 3), dtype=torch.float64, device=dev)
        last = self.end_idx
        force[last] = self.end_force * self.xs[last]
        self.force = force
        Xe, fe, we = self.X[last], force[last], self.weights[last]
        pp = self.pair_precision_rhs = "f64"
        vff = pairsum("oseen", Xe, fe, self.X, src_w=we, excl=self.excl_end,
                      eta=self.eta, reg=self.reg, eps=self.eps, precision=pp)
        self.v_explicit = vff  # explicit flow on fiber nodes (feature of the learned guess)
        rhs = torch.zeros(self.size, dtype=torch.float64, device=dev)
        B = self.bodies
        if B is not None:
            # external body force/torque: Stokeslet + rotlet at each centre
            B.explicit(self.X, out=vff, accumulate=True)
            vb = pairsum("oseen", Xe, fe, B.pos, src_w=we, eta=self.eta,
                         reg=self.reg, eps=self.eps, precision=pp)
            B.explicit(B.pos, out=vb, accumulate=True)
            B.scatter_density(rhs, -vb, 3 * self.nper)
        if self.shell is not None:
            vfs = pairsum("oseen", Xe, fe, self.shell.pos, src_w=we,
                          eta=self.eta, reg=self.reg, eps=self.eps, precision=pp)
            if B is not None:
                B.explicit(self.shell.pos, out=vfs, accumulate=True)
            rhs[:3 * self.nper] = -vfs.reshape(-1)
        dt, c0, c1, p = self.dt, self.c0, self.c1, self.penalty
        if self.assembly == "hip":
            # fused per-fiber RHS rows, one launch per node count
            lib = L.load()
            for N, bk in self.buckets.items():
                t = bk["asm"]
                rc = lib.f3d_fiber_rhs(
                    len(bk["ids"]), N, L.ptr(t["node0"]), L.ptr(t["length"]), L.ptr(t["rhs_off"]),
                    L.ptr(self.X), L.ptr(self.xs), L.ptr(self.xss), L.ptr(vff), L.ptr(t["D0"]),
                    L.ptr(t["PX"]), L.ptr(t["PT"]), float(dt), float(c0), float(c1), float(p),
                    float(self.end_force), L.ptr(rhs), _stream())
                L.check(rc, "f3d_fiber_rhs")
            self.rhs = rhs
            return
        for N, (node, X, xs, xss, D1, P, B) in self._ctx.items():
            bk = self.buckets[N]
            fk = force[node]  # (B, N, 3)
            vk = vff[node]
            alpha = bk["alpha"]
            R = torch.zeros((B, 4 * N), dtype=torch.float64, device=dev)
            xf = (xs * fk).sum(-1)
            for a in range(3):
                R[:, a * N:(a + 1) * N] = X[..., a] / dt + vk[..., a] \
                    + c0 * (fk[..., a] + xs[..., a] * xf) + c1 * (fk[..., a] - xs[..., a] * xf)
            R[:, 3 * N:] = -p + (xs * torch.matmul(D1, vk)).sum(-1) \
                + 2.0 * c0 * (xs * torch.matmul(D1, fk)).sum(-1) + (c0 - c1) * (xss * fk).sum(-1)
            Rsq = torch.zeros((B, 4 * N), dtype=torch.float64, device=dev)
            Rsq[:, :4 * N - 14] = torch.matmul(P, R[..., None])[..., 0]
            Fend = fk[:, -1]
            Rsq[:, 4 * N - 7:4 * N - 4] = Fend
            Rsq[:, 4 * N - 4] = (Fend * xs[:, -1]).sum(-1)
            _scatter_blocks(rhs, self.fib0 + 4 * self.offs[bk["ids"]], Rsq)
        self.rhs = rhs

    # ------------------------------------------------------------------
    def matvec(self, x, out=None, precision=None):
        """A_body_fiber_hyd (tstep.py:1003-1178) for fibers, free bodies and shell."""
        prec = precision or self.pair_precision
        if out is None:
            out = torch.empty_like(x)
        nper3 = 3 * self.nper
        XT = x[self.fib0:]
        B = self.bodies
        kw = dict(eta=self.eta, reg=self.reg, eps=self.eps, precision=prec)
        # fiber forces f = F XT into SoA (3, npts)
        self.gF(self.F_buf, XT, self.f_soa)
        fsoa = self._soa(self.f_soa)
        vsoa = self._soa(self.v_soa)
        pairsum("oseen", self.X, fsoa, self.X, src_w=self.weights, excl=self.excl, out=vsoa, **kw)
        ft = self._link_forces(XT) if self.att else None

        def link_flow(trg, out):
            # Stokeslet + rotlet of the attached fibers' forces and torques at
            # the body centres (tstep.py:1052-1064)
            pairsum("oseen", B.centres, ft[:, :3].contiguous(), trg, out=out, accumulate=True, **kw)
            pairsum("rotlet", B.centres, ft[:, 3:].contiguous(), trg, out=out, accumulate=True,
                    **kw)
        if B is not None:
            lam = B.density_view(x, nper3)
            vb = B.vb
            # body double layer on fibers and bodies (the j = i pair dropped)
            pairsum("stresslet", B.pos, lam, self.X, src_nrm=B.nrm, out=vsoa, accumulate=True, **kw)
            pairsum("stresslet", B.pos, lam, B.pos, src_nrm=B.nrm, out=vb, exclude_self=True, **kw)
            pairsum("oseen", self.X, fsoa, B.pos, src_w=self.weights, out=vb, accumulate=True, **kw)
            if ft is not None:
                link_flow(self.X, vsoa)
                link_flow(B.pos, vb)
        if self.shell is not None:
            q = x[:nper3].view(self.nper, 3)
            pairsum("oseen", self.X, fsoa, self.shell.pos, src_w=self.weights, out=self.v_shell, **kw)
            if B is not None:
                pairsum("stresslet", B.pos, lam, self.shell.pos, src_nrm=B.nrm, out=self.v_shell,
                        accumulate=True, **kw)
                pairsum("stresslet", self.shell.pos, q, B.pos, src_nrm=self.shell.nrm, out=vb,
                        accumulate=True, **kw)
                if ft is not None:
                    link_flow(self.shell.pos, self.v_shell)
            self.shell.apply(q, add=self.v_shell, out=out[:nper3].view(self.nper, 3),
                             precision=prec)
            pairsum("stresslet", self.shell.pos, q, self.X, src_nrm=self.shell.nrm, out=vsoa,
                    accumulate=True, **kw)
        if B is not None:
            B.rows(x, out, nper3)
        res = out[self.fib0:]
        self.gA(self.A_buf, XT, res)
        for c in range(3):
            self.gQ[c](self.Q_buf[c], self.v_soa, res, alpha=-1.0, beta=1.0)
        if self.att:
            self._link_rows(x, out, vsoa)
        # minus ends clamped to the cortex: flow term of their tension
        # condition (tstep.py:1162-1166)
        for N, g in self.cor.items():
            out.index_add_(0, g["row0"] + 3, (vsoa[g["node0"]] * g["xs0"]).sum(1))
        # hinged plus ends: flow term of their tension condition (tstep.py:1167-1168)
        for N, g in self.hng.items():
            out.index_add_(0, g["row0"] + 3, (vsoa[g["nodeL"]] * g["xsL"]).sum(1))
        return out

    def precond(self, x, out=None):
        """P_inv (tstep.py:950-990)."""
        if out is None:
            out = torch.empty_like(x)
        nper3 = 3 * self.nper
        if self.shell is not None:
            self.shell.precond(x[:nper3], out[:nper3], f32=self.lu_precision == "f32c")
        if self.bodies is not None:
            self.bodies.precond(x, out, nper3)
        self.lu.solve(x[self.fib0:], out=out[self.fib0:])
        return out

    def dense_blocks_host(self):
        """Host copies of the per-fiber A_sq blocks (for checks)."""
        return self.bs.unpack(self.A_buf)
