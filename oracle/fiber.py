"""Oracle: per-fiber discretisation and implicit operator (FP64, NumPy).

Restates (paths relative to the reference's python/ directory):
  utils/finite_diff.py:4-63          finite_diff (Fornberg 1998 weights)
  utils/barycentricMatrix.py:12-38   barycentric_matrix
  utils/cheb.py:4-29, :50-75         cheb, clencurt (Chebyshev variant)
  tstep/fiber_matrices.py:38-153     FiberMatrices (uniform nodes + FD: 'reference')
  fiber/fiber.py:85-87               c0, c1
  fiber/fiber.py:322-447             form_linear_operator (penalty formulation)
  fiber/fiber.py:473-582             compute_RHS
  fiber/fiber.py:810-961             apply_BC_rectangular (force/torque ends)
  fiber/fiber.py:1399-1431           force_operator
  tstep/tstep.py:700-711             derivatives of the configuration

Pinned by tests/test_oracle_fiber.py: FD weights are exact on polynomials of
degree below the stencil size (and D 1 = 0); the analytic-derivative check of
tests/test_finite_diff.py:50-54; barycentric interpolation reproduces
constants and is the identity on its own nodes; and golden vectors written by
tests/golden/make_fiber_golden.py from the reference's own fiber.py /
fiber_matrices.py (importable here: NumPy/SciPy only) pin the assembly.

The 'chebyshev' kind (Chebyshev-Lobatto nodes, spectral D, first-kind
collocation grids, Clenshaw-Curtis weights) restates, as read here, the
older generation (utils/cheb.py, multi_fibers/multi_fiber.py:97); it has no
live reference counterpart: parity unpinned beyond the polynomial checks.
"""
import numpy as np


# ---------------------------------------------------------------------------
def finite_diff(s, M, n_s):
    """Fornberg weights: D[i, :, m] is the m-th derivative row at s[i], m <= M,
    on a stencil of n_s points, one-sided near the ends (finite_diff.py:4-63)."""
    s = np.asarray(s, dtype=np.float64)
    npts = s.size
    D = np.zeros((npts, npts, M + 1))
    half = (n_s - 1) // 2
    last = n_s - 1
    for xi, si in enumerate(s):
        if xi < half:
            lo = 0
        elif xi > npts - half - 2:
            lo = npts - last - 1
        else:
            lo = xi - half
        x = s[lo:lo + last + 1]
        c = np.zeros((last + 1, M + 1))
        c[0, 0] = 1.0
        c1 = 1.0
        c4 = x[0] - si
        for i in range(1, last + 1):
            mn = min(i, M)
            c2 = 1.0
            c5 = c4
            c4 = x[i] - si
            for j in range(i):
                c3 = x[i] - x[j]
                c2 *= c3
                if j == i - 1:
                    for k in range(mn, 0, -1):
                        c[i, k] = c1 * (k * c[i - 1, k - 1] - c5 * c[i - 1, k]) / c2
                    c[i, 0] = -c1 * c5 * c[i - 1, 0] / c2
                for k in range(mn, 0, -1):
                    c[j, k] = (c4 * c[j, k] - k * c[j, k - 1]) / c3
                c[j, 0] = c4 * c[j, 0] / c3
            c1 = c2
        D[xi, lo:lo + last + 1, :] = c
    return D


def barycentric_matrix(x, y):
    """Resampling matrix P (len(y) x len(x)) with weights (-1)^k, halved at both
    ends (barycentricMatrix.py:12-38); rows hitting a node copy that value."""
    x = np.asarray(x, float).ravel()
    y = np.asarray(y, float).ravel()
    N = x.size
    w = np.ones(N)
    w[1::2] = -1.0
    w[0] = 0.5
    w[-1] = -0.5 * (-1) ** N
    with np.errstate(divide="ignore", invalid="ignore"):
        yx = y[:, None] - x
        wyx = w[None, :] / yx
        P = wyx / np.sum(wyx, axis=1)[:, None]
    P[yx == 0] = 1.0
    return P


def cheb(N):
    """Chebyshev differentiation matrix and extrema cos(pi s/N), s=0..N (cheb.py:4-29)."""
    if N == 0:
        return np.zeros((1, 1)), np.ones(1)
    s = np.arange(N + 1, dtype=float)
    x = np.cos(np.pi * s / N)
    c = np.hstack(([2.0], np.ones(N - 1), [2.0])) * (-1.0) ** s
    dX = x[:, None] - x[None, :]
    D = np.outer(c, 1.0 / c) / (dX + np.eye(N + 1))
    D -= np.diag(D.sum(axis=1))
    return D, x


def clencurt(N):
    """Clenshaw-Curtis weights on the N+1 Chebyshev extrema (cheb.py:50-75)."""
    theta = np.pi / N * np.arange(N + 1)
    w = np.zeros(N + 1)
    ii = np.arange(1, N)
    v = np.ones(N - 1)
    if N % 2 == 0:
        w[0] = w[N] = 1.0 / (N * N - 1.0)
        for k in range(1, N // 2):
            v -= 2.0 * np.cos(2.0 * k * theta[ii]) / (4.0 * k * k - 1)
        v -= np.cos(N * theta[ii]) / (N * N - 1)
    else:
        w[0] = w[N] = 1.0 / N**2
        for k in range(1, (N - 1) // 2 + 1):
            v -= 2.0 * np.cos(2.0 * k * theta[ii]) / (4.0 * k * k - 1)
    w[ii] = 2.0 * v / N
    return w


class FiberMatrices:
    """Per-resolution fiber matrices (tstep/fiber_matrices.py:38-104)."""

    def __init__(self, num_points, nfd=4, kind="reference"):
        N = int(num_points)
        self.num_points = N
        self.kind = kind
        if kind == "reference":
            self.alpha = np.linspace(-1.0, 1.0, N)
            nr, nt = N - 4, N - 2
            self.alpha_roots = 2.0 * (0.5 + np.arange(nr)) / nr - 1.0
            self.alpha_tension = 2.0 * (0.5 + np.arange(nt)) / nt - 1.0
            self.D0 = [finite_diff(self.alpha, k, nfd + k)[:, :, k] for k in (1, 2, 3, 4)]
            w = np.full(N, 2.0)
            w[0] = w[-1] = 1.0
            self.weights0 = w / (N - 1)
        elif kind == "chebyshev":
            Dc, xc = cheb(N - 1)
            self.alpha = xc[::-1].copy()
            D1 = Dc[::-1, ::-1].copy()
            nr, nt = N - 4, N - 2
            self.alpha_roots = -np.cos(np.pi * (2 * np.arange(nr) + 1) / (2 * nr))
            self.alpha_tension = -np.cos(np.pi * (2 * np.arange(nt) + 1) / (2 * nt))
            self.D0 = [D1, D1 @ D1, D1 @ D1 @ D1, D1 @ D1 @ D1 @ D1]
            self.weights0 = clencurt(N - 1)[::-1].copy()
        else:
            raise ValueError(kind)
        self.P_X = barycentric_matrix(self.alpha, self.alpha_roots)
        self.P_T = barycentric_matrix(self.alpha, self.alpha_tension)
        P = np.zeros((4 * N - 14, 4 * N))
        P[0:N - 4, 0:N] = self.P_X
        P[N - 4:2 * N - 8, N:2 * N] = self.P_X
        P[2 * N - 8:3 * N - 12, 2 * N:3 * N] = self.P_X
        P[3 * N - 12:4 * N - 14, 3 * N:4 * N] = self.P_T
        self.P_cheb = P

    def Ds(self, length):
        """D_k scaled by (2/L)^k (fiber_matrices.py:122-126)."""
        return [D * (2.0 / length) ** (k + 1) for k, D in enumerate(self.D0)]

    def weights(self, length):
        """Quadrature weights scaled by L/2 (fiber_matrices.py:134-136)."""
        return self.weights0 * (length / 2.0)

    def s(self, length):
        return (1.0 + self.alpha) * (length / 2.0)


def constants(epsilon, viscosity):
    """c0 = -ln(e eps^2)/(8 pi eta), c1 = 2/(8 pi eta) (fiber.py:85-86)."""
    c0 = -np.log(np.e * epsilon**2) / (8.0 * np.pi * viscosity)
    c1 = 2.0 / (8.0 * np.pi * viscosity)
    return c0, c1


def derivatives(fm, x, length):
    """xs, xss, xsss, xssss = D_k x (tstep.py:700-711)."""
    return [D @ x for D in fm.Ds(length)]


def form_linear_operator(fm, x, length, E, dt, c0, c1, penalty=500.0, beta_tstep=1.0):
    """4N x 4N implicit operator (fiber.py:322-447, penalty, inonlocal=0)."""
    N = fm.num_points
    D1, D2, D3, D4 = fm.Ds(length)
    xs, xss, xsss, _ = derivatives(fm, x, length)
    I = np.eye(N)

    def dl(v, D):  # diag(v) @ D
        return v[:, None] * D

    A = np.zeros((4 * N, 4 * N))
    for a in range(3):
        for b in range(a, 3):
            if a == b:
                blk = beta_tstep * I / dt + E * c0 * dl(1.0 + xs[:, a] ** 2, D4) \
                    + E * c1 * dl(1.0 - xs[:, a] ** 2, D4)
            else:
                blk = E * (c0 - c1) * dl(xs[:, a] * xs[:, b], D4)
            A[a * N:(a + 1) * N, b * N:(b + 1) * N] = blk
            A[b * N:(b + 1) * N, a * N:(a + 1) * N] = blk
        A[a * N:(a + 1) * N, 3 * N:] = -c0 * (2.0 * dl(xs[:, a], D1) + np.diag(xss[:, a])) \
            - c1 * np.diag(xss[:, a])
        A[3 * N:, a * N:(a + 1) * N] = -(c1 + 7.0 * c0) * E * dl(xss[:, a], D4) \
            - 6.0 * c0 * E * dl(xsss[:, a], D3) - penalty * dl(xs[:, a], D1)
    A[3 * N:, 3 * N:] = -2.0 * c0 * D2 + (c1 + c0) * np.diag(np.sum(xss**2, axis=1))
    return A


def compute_RHS(fm, x, length, dt, c0, c1, penalty=500.0, force=None, flow=None, v_length=0.0):
    """RHS of the implicit system (fiber.py:473-582, inonlocal=0)."""
    N = fm.num_points
    D1 = fm.Ds(length)[0]
    xs, xss, _, _ = derivatives(fm, x, length)
    alpha = fm.alpha
    R = np.zeros(4 * N)
    for a in range(3):
        R[a * N:(a + 1) * N] = x[:, a] / dt + (alpha + 1.0) * (v_length * 0.5) * xs[:, a]
    R[3 * N:] = -penalty
    if flow is not None:
        for a in range(3):
            R[a * N:(a + 1) * N] += flow[:, a]
        R[3 * N:] += np.sum(xs * (D1 @ flow), axis=1)
    if force is not None:
        fs = D1 @ force
        xf = np.sum(xs * force, axis=1)
        for a in range(3):
            # c0 (I + xs xs^T) f + c1 (I - xs xs^T) f
            R[a * N:(a + 1) * N] += c0 * (force[:, a] + xs[:, a] * xf) \
                + c1 * (force[:, a] - xs[:, a] * xf)
        R[3 * N:] += 2.0 * c0 * np.sum(xs * fs, axis=1) + (c0 - c1) * np.sum(xss * force, axis=1)
    return R


def apply_BC_rectangular(fm, x, length, E, A, RHS, end_force=None,
                         start_force=None, start_torque=None, end_torque=None,
                         clamped=False, dt=None, c0=None, flow=None, force=None,
                         hinged=False):
    """Square system [P_cheb A; B], [P_cheb RHS; B_RHS] for force/torque-free
    ends with optional prescribed end forces/torques (fiber.py:810-961, the
    'force' and 'torque' branches).  clamped=True replaces the minus-end rows
    by the 'velocity' and 'angular_velocity' branches (fiber.py:855-877,
    :887-898) with zero prescribed velocities, as tstep sets them for a fiber
    attached to a body (tstep.py:878-881): x(0)/dt and x_s(0)/dt rows, and
    the tension row 6 E c0 x_ss(0) . D3 X + 2 c0 D1 T whose right-hand side
    carries -(flow(0) + 2 c0 force(0)) . x_s(0).  hinged=True replaces the
    plus-end force rows by the 'velocity' branch at s = L (fiber.py:940-960),
    as for a fiber that reached the periphery (tstep.py:886-888): x(L)/dt and
    the same tension row at L; the torque rows stay."""
    N = fm.num_points
    D1, D2, D3, _ = fm.Ds(length)
    xs, xss, _, _ = derivatives(fm, x, length)
    z3 = np.zeros(3)
    F0 = z3 if start_force is None else np.asarray(start_force, float)
    L0 = z3 if start_torque is None else np.asarray(start_torque, float)
    F1 = z3 if end_force is None else np.asarray(end_force, float)
    L1 = z3 if end_torque is None else np.asarray(end_torque, float)
    B = np.zeros((14, 4 * N))
    Bv = np.zeros(14)
    for a in range(3):
        B[a, a * N:(a + 1) * N] = E * D3[0]
        B[a, 3 * N] = -xs[0, a]
        B[3, a * N:(a + 1) * N] = -E * D2[0] * xss[0, a]
        B[4 + a, a * N:(a + 1) * N] = D2[0]
        B[7 + a, a * N:(a + 1) * N] = -E * D3[-1]
        B[7 + a, 4 * N - 1] = xs[-1, a]
        B[10, a * N:(a + 1) * N] = E * D2[-1] * xss[-1, a]
        B[11 + a, a * N:(a + 1) * N] = D2[-1]
    B[3, 3 * N] = -1.0
    B[10, 4 * N - 1] = 1.0
    Bv[0:3] = F0
    Bv[3] = np.dot(F0, xs[0])
    Bv[4:7] = L0
    Bv[7:10] = F1
    Bv[10] = np.dot(F1, xs[-1])
    Bv[11:14] = L1
    if clamped:
        B[0:7] = 0.0
        for a in range(3):
            B[a, a * N] = 1.0 / dt
            B[3, a * N:(a + 1) * N] = 6.0 * E * c0 * xss[0, a] * D3[0]
            B[4 + a, a * N:(a + 1) * N] = D1[0] / dt
        B[3, 3 * N:] = 2.0 * c0 * D1[0]
        Bv[0:3] = x[0] / dt
        Bv[3] = 0.0
        if flow is not None:
            Bv[3] -= np.dot(np.asarray(flow, float)[0], xs[0])
        if force is not None:
            Bv[3] -= 2.0 * c0 * np.dot(np.asarray(force, float)[0], xs[0])
        Bv[4:7] = xs[0] / dt
    if hinged:
        B[7:11] = 0.0
        for a in range(3):
            B[7 + a, a * N + N - 1] = 1.0 / dt
            B[10, a * N:(a + 1) * N] = 6.0 * E * c0 * xss[-1, a] * D3[-1]
        B[10, 3 * N:] = 2.0 * c0 * D1[-1]
        Bv[7:10] = x[-1] / dt
        Bv[10] = 0.0
        if flow is not None:
            Bv[10] -= np.dot(np.asarray(flow, float)[-1], xs[-1])
        if force is not None:
            Bv[10] -= 2.0 * c0 * np.dot(np.asarray(force, float)[-1], xs[-1])
    A_sq = np.vstack((fm.P_cheb @ A, B))
    b = np.concatenate((fm.P_cheb @ RHS, Bv))
    return A_sq, b


def force_operator(fm, x, length, E):
    """3N x 4N map XT -> f = -E X'''' + (T X_s)_s (fiber.py:1399-1431)."""
    N = fm.num_points
    D1, _, _, D4 = fm.Ds(length)
    xs, xss, _, _ = derivatives(fm, x, length)
    F = np.zeros((3 * N, 4 * N))
    for a in range(3):
        F[a * N:(a + 1) * N, a * N:(a + 1) * N] = -E * D4
        F[a * N:(a + 1) * N, 3 * N:] = np.diag(xss[:, a]) + xs[:, a][:, None] * D1
    return F
