"""Oracle: confining-periphery (shell) double-layer operator (FP64, NumPy).

Two independent restatements of the same operator, from two places in the
reference, so that each checks the other:
  * shell_matrix: the dense 3N x 3N matrix built at tstep/tstep.py:328-345
    (stresslet . n matrix, minus the singularity-subtraction block, minus
    diag(1/w), plus the complementary n n^T), applied at tstep.py:1099-1101;
  * shell_apply: the matrix-free form of multi_bodies/multi_bodies.py:341-381.
singularity_subtraction_vectors follows periphery/periphery.py:75-98.
"""
import numpy as np

from . import kernels as K


def singularity_subtraction_vectors(x, n, w, eta=1.0, reg=5e-3, eps=1e-5):
    """ex, ey, ez (each (N,3)): stresslet self-sums of w * e_c (periphery.py:75-98)."""
    x = np.asarray(x, float).reshape(-1, 3)
    N = x.shape[0]
    out = []
    for c in range(3):
        e = np.zeros((N, 3))
        e[:, c] = 1.0
        e *= np.asarray(w, float).ravel()[:, None]
        out.append(K.stresslet_kernel_times_normal_times_density(x, n, e, eta, reg, eps).reshape(N, 3))
    return out


def shell_matrix(x, n, w, ex, ey, ez, eta=1.0, reg=5e-3, eps=1e-5):
    """Dense shell operator M (tstep.py:328-345), AoS ordering."""
    x = np.asarray(x, float).reshape(-1, 3)
    N = x.shape[0]
    w = np.asarray(w, float).ravel()
    M = K.stresslet_kernel_times_normal(x, n, eta, reg, eps)
    I = np.zeros((3 * N, 3 * N))
    for i in range(N):
        I[3 * i:3 * (i + 1), 3 * i + 0] = ex[i] / w[i]
        I[3 * i:3 * (i + 1), 3 * i + 1] = ey[i] / w[i]
        I[3 * i:3 * (i + 1), 3 * i + 2] = ez[i] / w[i]
    I_vec = np.repeat(1.0 / w, 3)
    M += -I - np.diag(I_vec)
    M += K.complementary_kernel(x, n)
    return M


def shell_apply(x, n, w, ex, ey, ez, q, eta=1.0, reg=5e-3, eps=1e-5):
    """Matrix-free shell operator (multi_bodies/multi_bodies.py:341-381)."""
    x = np.asarray(x, float).reshape(-1, 3)
    N = x.shape[0]
    v = np.asarray(q, float).reshape(N, 3)
    w = np.asarray(w, float).ravel()
    res = K.stresslet_kernel_times_normal_times_density(x, n, v, eta, reg, eps)
    res += K.complementary_kernel_times_density(x, n, v)
    cx = ((v[:, 0] / w)[:, None] * ex).ravel()
    cy = ((v[:, 1] / w)[:, None] * ey).ravel()
    cz = ((v[:, 2] / w)[:, None] * ez).ravel()
    res -= cx + cy + cz
    res -= (v / w[:, None]).ravel()
    return res
