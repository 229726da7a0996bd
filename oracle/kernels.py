"""Oracle: Green's-function pair sums and dense kernel matrices (FP64, NumPy).

Restates python/kernels/kernels.py.  All inputs are float64 arrays of 3N
values, flat or (N,3), AoS ("xyzxyz") as in the reference; outputs are flat.

Pinned by tests/test_oracle_kernels.py:
  * Oseen: Stokes drag identity (mean of the Stokeslet over a sphere of
    radius R around the source is F/(6 pi eta R)), divergence-free field,
    reciprocity of the symmetric tensor, regularised coincident pair equals
    f/(8 pi eta reg) as the reference's branch implies.
  * Rotlet: equals half the curl of the Oseen field (finite differences).
  * Stresslet: double-layer Gauss identity on a Fibonacci sphere
    (-e inside, 0 outside for this sign convention and outward normals) and
    divergence-free field; self-sum j != i equals the dense matrix apply.
  * Complementary kernel: rank-one algebra.
"""
import numpy as np

_CHUNK = 2048  # targets per vectorised block


def _pts(a):
    a = np.asarray(a, dtype=np.float64)
    return a.reshape(a.size // 3, 3)


def _dist(d, eps, reg):
    """r = |d|, regularised to sqrt(r^2+reg^2) where r < eps (kernels.py:830-831)."""
    r = np.sqrt(np.einsum("...k,...k->...", d, d))
    near = r < eps
    if np.any(near):
        r = np.where(near, np.sqrt(r * r + reg * reg), r)
    return r


def oseen_kernel_source_target(r_source, r_target, density, eta=1.0, reg=5e-3,
                               epsilon_distance=1e-5, exclude=None):
    """u_i = sum_j 1/(8 pi eta) [f_j/r + d (d.f_j)/r^3], d = t_i - s_j.

    Follows kernels.oseen_kernel_source_target_numba (kernels/kernels.py:794-846).
    exclude: optional (lo, hi) int arrays; pairs with lo[i] <= j < hi[i] are
    dropped (the GPU path's same-fiber exclusion).
    """
    s, t, f = _pts(r_source), _pts(r_target), _pts(density)
    factor = 1.0 / (8.0 * np.pi * eta)
    u = np.zeros_like(t)
    for i0 in range(0, t.shape[0], _CHUNK):
        tt = t[i0:i0 + _CHUNK]
        d = tt[:, None, :] - s[None, :, :]
        r = _dist(d, epsilon_distance, reg)
        fr = factor / r
        gr = factor / r**3
        if exclude is not None:
            lo, hi = exclude
            j = np.arange(s.shape[0])[None, :]
            m = (j >= lo[i0:i0 + _CHUNK, None]) & (j < hi[i0:i0 + _CHUNK, None])
            fr = np.where(m, 0.0, fr)
            gr = np.where(m, 0.0, gr)
        df = np.einsum("ijk,jk->ij", d, f)
        u[i0:i0 + _CHUNK] = fr @ f + np.einsum("ij,ijk->ik", gr * df, d)
    return u.ravel()


def rotlet_kernel_source_target(r_source, r_target, density, eta=1.0, reg=5e-3,
                                epsilon_distance=1e-5, exclude=None):
    """u_i = sum_j (tau_j x d) / (8 pi eta r^3).

    Follows kernels.rotlet_kernel_source_target_numba (kernels/kernels.py:856-908):
    u_x += fr (z tau_y - y tau_z), etc., fr = 1/(8 pi eta r^3).
    """
    s, t, tau = _pts(r_source), _pts(r_target), _pts(density)
    factor = 1.0 / (8.0 * np.pi * eta)
    u = np.zeros_like(t)
    for i0 in range(0, t.shape[0], _CHUNK):
        tt = t[i0:i0 + _CHUNK]
        d = tt[:, None, :] - s[None, :, :]
        r = _dist(d, epsilon_distance, reg)
        fr = factor / r**3
        if exclude is not None:
            lo, hi = exclude
            j = np.arange(s.shape[0])[None, :]
            fr = np.where((j >= lo[i0:i0 + _CHUNK, None]) & (j < hi[i0:i0 + _CHUNK, None]), 0.0, fr)
        cr = np.cross(np.broadcast_to(tau[None], d.shape), d)
        u[i0:i0 + _CHUNK] = np.einsum("ij,ijk->ik", fr, cr)
    return u.ravel()


def stresslet_kernel_source_target(r_source, r_target, normal, density, eta=1.0, reg=5e-3,
                                   epsilon_distance=1e-5, exclude=None):
    """u_i = sum_j (-3/(4 pi eta)) (d.q_j)(d.n_j) d / r^5.

    Follows kernels.stresslet_kernel_source_target_numba (kernels/kernels.py:1186-1217).
    """
    s, t = _pts(r_source), _pts(r_target)
    n, q = _pts(normal), _pts(density)
    factor = -3.0 / (4.0 * np.pi * eta)
    u = np.zeros_like(t)
    for i0 in range(0, t.shape[0], _CHUNK):
        tt = t[i0:i0 + _CHUNK]
        d = tt[:, None, :] - s[None, :, :]
        r = _dist(d, epsilon_distance, reg)
        f0 = factor * np.einsum("ijk,jk->ij", d, q) * np.einsum("ijk,jk->ij", d, n) / r**5
        if exclude is not None:
            lo, hi = exclude
            j = np.arange(s.shape[0])[None, :]
            f0 = np.where((j >= lo[i0:i0 + _CHUNK, None]) & (j < hi[i0:i0 + _CHUNK, None]), 0.0, f0)
        u[i0:i0 + _CHUNK] = np.einsum("ij,ijk->ik", f0, d)
    return u.ravel()


def stresslet_kernel_times_normal_times_density(r_vectors, normal, density, eta=1.0, reg=5e-3,
                                                epsilon_distance=1e-5):
    """Self-sum over j != i of the stresslet contracted with normal and density.

    Follows kernels.stresslet_kernel_times_normal_times_density_numba
    (kernels/kernels.py:1123-1183): the loop skips yn == xn explicitly.
    """
    x = _pts(r_vectors)
    N = x.shape[0]
    idx = np.arange(N, dtype=np.int64)
    return stresslet_kernel_source_target(x, x, normal, density, eta, reg, epsilon_distance,
                                          exclude=(idx, idx + 1))


def stresslet_kernel_times_normal(r_vectors, normal, eta=1.0, reg=5e-3, epsilon_distance=1e-5):
    """Dense 3N x 3N matrix, block (i, j != i) = (-3/(4 pi eta)) (d.n_j)/r^5 d d^T, d = x_i - x_j.

    Follows kernels.stresslet_kernel_times_normal_numba (kernels/kernels.py:1063-1120).
    """
    x, n = _pts(r_vectors), _pts(normal)
    N = x.shape[0]
    factor = -3.0 / (4.0 * np.pi * eta)
    d = x[:, None, :] - x[None, :, :]
    r = _dist(d, epsilon_distance, reg)
    np.fill_diagonal(r, 1.0)
    c = factor * np.einsum("ijk,jk->ij", d, n) / r**5
    np.fill_diagonal(c, 0.0)
    S = c[:, :, None, None] * d[:, :, :, None] * d[:, :, None, :]
    return S.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)


def complementary_kernel(r_vectors, normal):
    """Nk = n n^T on the flattened normal vector (kernels/kernels.py:1286-1295)."""
    n = np.asarray(normal, dtype=np.float64).ravel()
    return np.outer(n, n)


def complementary_kernel_times_density(r_vectors, normal, density):
    """(Nk q)_i = n_i sum_j n_j . q_j (kernels/kernels.py:1267-1283)."""
    n, q = _pts(normal), _pts(density)
    return (n * np.sum(n * q)).ravel()


def _xyz_points(r, fmt):
    """(N, 3) points from AoS ('r') or concatenated x, y, z blocks ('xyz')."""
    r = np.asarray(r, dtype=np.float64).ravel()
    if fmt == "r":
        return r.reshape(-1, 3)
    return np.ascontiguousarray(r.reshape(3, -1).T)


def oseen_tensor_source_target(r_source, r_target, eta=1.0, reg=5e-3, epsilon_distance=1e-5,
                               input_format="r"):
    """Dense 3Nt x 3Ns Oseen tensor (kernels/kernels.py:119-211): block (i, j)
    = (I/r + d d^T/r^3)/(8 pi eta), d = t_i - s_j; entries with r <= eps
    (note <=, unlike the numba sums) use sqrt(r^2 + reg^2).  'r' orders rows
    and columns point by point, 'xyz' component block by component block."""
    s, t = _xyz_points(r_source, input_format), _xyz_points(r_target, input_format)
    Ns, Nt = s.shape[0], t.shape[0]
    d = t[:, None, :] - s[None, :, :]
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    r = np.where(r > epsilon_distance, r, np.sqrt(r * r + reg * reg))
    fr = 1.0 / (8.0 * np.pi * eta * r)
    gr = 1.0 / (8.0 * np.pi * eta * r**3)
    G = fr[:, :, None, None] * np.eye(3)[None, None] + gr[:, :, None, None] * d[:, :, :, None] * d[:, :, None, :]
    if input_format == "r":
        return G.transpose(0, 2, 1, 3).reshape(3 * Nt, 3 * Ns)
    return G.transpose(2, 0, 3, 1).reshape(3 * Nt, 3 * Ns)


def oseen_tensor(r_vectors, eta=1.0, reg=5e-3, epsilon_distance=1e-5, input_format="r"):
    """Dense 3N x 3N Oseen tensor of N points (kernels/kernels.py:31-116): the
    source-target tensor with the points as both; the diagonal block is
    I/(8 pi eta reg)."""
    return oseen_tensor_source_target(r_vectors, r_vectors, eta, reg, epsilon_distance,
                                      input_format)
