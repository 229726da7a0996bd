"""Per-resolution fiber matrices (host, one-off per node count; cached).

Behaviour of tstep/fiber_matrices.py:38-153 (reference discretisation: uniform
parameter nodes alpha = linspace(-1, 1, N), finite differences of stencil
nfd + k for the k-th derivative, one-sided near the ends (weights solved
exactly in rational arithmetic on the integer stencil offsets), uniform-midpoint
collocation grids of N-4 and N-2 points, trapezoid weights), plus a
Chebyshev-Lobatto variant.  The FD weights come from solving the scaled
Vandermonde system of each stencil (exact for polynomials below the stencil
size), which yields the same weights as the reference's Fornberg recursion
(utils/finite_diff.py:4-63) to rounding.
"""
from fractions import Fraction
from functools import lru_cache
from math import factorial

import numpy as np


@lru_cache(maxsize=None)
def _exact_weights(offsets, k):
    """Weights c_j with sum_j c_j f(j) = f^(k)(0) + O(h^len) on integer offsets,
    solved exactly in rationals (Vandermonde system sum_j c_j j^p = k! [p == k])."""
    n = len(offsets)
    A = [[Fraction(o) ** p for o in offsets] + [Fraction(factorial(k) if p == k else 0)]
         for p in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if A[r][col] != 0)
        A[col], A[piv] = A[piv], A[col]
        for r in range(n):
            if r != col and A[r][col] != 0:
                f = A[r][col] / A[col][col]
                A[r] = [a - f * b for a, b in zip(A[r], A[col])]
    return tuple(A[r][n] / A[r][r] for r in range(n))


def fd_matrix_uniform(n, k, npts):
    """Order-k FD matrix on n uniform nodes of [-1, 1] (spacing 2/(n-1)) with
    npts-point stencils, exact rational weights rounded once to FP64."""
    half = (npts - 1) // 2
    D = np.zeros((n, n))
    h = 2.0 / (n - 1)
    for i in range(n):
        if i < half:
            lo = 0
        elif i > n - half - 2:
            lo = n - npts
        else:
            lo = i - half
        offs = tuple(range(lo - i, lo - i + npts))
        D[i, lo:lo + npts] = [float(c) for c in _exact_weights(offs, k)]
        D[i, lo:lo + npts] /= h**k
    return D


def fd_matrix(alpha, k, npts):
    """Differentiation matrix of order k on nodes alpha with npts-point stencils."""
    alpha = np.asarray(alpha, float)
    n = alpha.size
    half = (npts - 1) // 2
    D = np.zeros((n, n))
    for i in range(n):
        lo = min(max(i - half, 0), n - npts)
        # the reference keeps the centred stencil until xi > n - half - 2
        if i >= half and i <= n - half - 2:
            lo = i - half
        idx = np.arange(lo, lo + npts)
        h = alpha[idx] - alpha[i]
        scale = np.max(np.abs(h))
        t = h / scale
        V = np.vander(t, npts, increasing=True).T  # V[p, j] = t_j^p
        rhs = np.zeros(npts)
        rhs[k] = factorial(k)
        c = np.linalg.solve(V, rhs)
        D[i, idx] = c / scale**k
    return D


def bary_matrix(x, y):
    """Barycentric resampling from nodes x to points y with the Chebyshev-Lobatto
    weights (+-1, halved at the ends); a point equal to a node copies it."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    w = (-1.0) ** np.arange(x.size)
    w[0] *= 0.5
    w[-1] = 0.5 * (-1.0) ** (x.size - 1)
    P = np.empty((y.size, x.size))
    for r, yr in enumerate(y):
        d = yr - x
        hit = np.flatnonzero(d == 0.0)
        if hit.size:
            P[r] = 0.0
            P[r, hit[0]] = 1.0
            continue
        t = w / d
        P[r] = t / t.sum()
    return P


def _cheb_D(N):
    """Spectral differentiation on ascending Chebyshev-Lobatto nodes."""
    m = N - 1
    x = -np.cos(np.pi * np.arange(N) / m)
    c = np.ones(N)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(N)
    D = np.outer(c, 1.0 / c) / (x[:, None] - x[None, :] + np.eye(N))
    D -= np.diag(D.sum(axis=1))
    return x, D


def _clenshaw_curtis(N):
    m = N - 1
    theta = np.pi * np.arange(N) / m
    w = np.zeros(N)
    for j in range(N):
        s = 0.0
        for k in range(1, m // 2 + 1):
            b = 1.0 if 2 * k == m else 2.0
            s += b * np.cos(2 * k * theta[j]) / (4 * k * k - 1)
        cj = 1.0 if j in (0, m) else 2.0
        w[j] = cj / m * (1.0 - s)
    return w


class FiberMatrices:
    def __init__(self, N, nfd=4, kind="reference"):
        self.N = N = int(N)
        self.kind = kind
        if kind == "reference":
            # the k-th derivative uses nfd + k of the N nodes
            # (num_points_finite_diff, fiber_matrices.py:24-31, :58-66); the
            # reference's Fornberg loop fails outright outside this range
            if not 1 <= nfd <= N - 4:
                raise ValueError("num_points_finite_diff must be in [1, N - 4] = [1, %d], got %r"
                                 % (N - 4, nfd))
            self.alpha = np.linspace(-1.0, 1.0, N)
            self.alpha_roots = -1.0 + (2.0 * np.arange(N - 4) + 1.0) / (N - 4)
            self.alpha_tension = -1.0 + (2.0 * np.arange(N - 2) + 1.0) / (N - 2)
            self.D0 = [fd_matrix_uniform(N, k, nfd + k) for k in (1, 2, 3, 4)]
            w = np.full(N, 2.0 / (N - 1))
            w[0] = w[-1] = 1.0 / (N - 1)
            self.weights0 = w
        elif kind == "chebyshev":
            self.alpha, D1 = _cheb_D(N)
            self.alpha_roots = -np.cos(np.pi * (2 * np.arange(N - 4) + 1) / (2 * (N - 4)))
            self.alpha_tension = -np.cos(np.pi * (2 * np.arange(N - 2) + 1) / (2 * (N - 2)))
            self.D0 = [np.linalg.matrix_power(D1, k) for k in (1, 2, 3, 4)]
            self.weights0 = _clenshaw_curtis(N)
        else:
            raise ValueError(kind)
        PX = bary_matrix(self.alpha, self.alpha_roots)
        PT = bary_matrix(self.alpha, self.alpha_tension)
        P = np.zeros((4 * N - 14, 4 * N))
        for c in range(3):
            P[c * (N - 4):(c + 1) * (N - 4), c * N:(c + 1) * N] = PX
        P[3 * (N - 4):, 3 * N:] = PT
        self.P_cheb = P


@lru_cache(maxsize=None)
def get(N, nfd=4, kind="reference"):
    return FiberMatrices(N, nfd, kind)
