"""Pins of the fiber-discretisation oracle.

* Golden vectors written by tests/golden/make_fiber_golden.py from the
  reference's own fiber.py / fiber_matrices.py pin the exact assembly of the
  operator, RHS, boundary rows and force operator.
* Polynomial exactness, D 1 = 0 and the analytic-derivative check of the
  reference's tests/test_finite_diff.py:50-54 pin the FD weights.
* Barycentric interpolation reproduces constants, is the identity on its own
  nodes, and interpolates the Gaussian of tests/test_barycentric.py:19-20.
"""
import os

import numpy as np
import pytest

from conftest import rel2
from oracle import fiber as OF

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fiber_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)  # arrays only; allow_pickle stays False


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_fiber_assembly_matches_reference_golden(gold, ci):
    g = {k[3:]: gold[k] for k in gold.files if k.startswith(f"c{ci}_")}
    N, L, E, dt = int(g["N"]), float(g["L"]), float(g["E"]), float(g["dt"])
    c0, c1 = OF.constants(float(g["eps"]), float(g["eta"]))
    fm = OF.FiberMatrices(N)
    x = g["x"]
    for k, D in enumerate(fm.Ds(L)):
        assert rel2(D, g[f"D{k + 1}"]) < 1e-13
    assert rel2(fm.weights(L), g["weights"]) < 1e-15
    assert rel2(fm.P_cheb, g["P_cheb"]) < 1e-14
    A = OF.form_linear_operator(fm, x, L, E, dt, c0, c1)
    assert rel2(A, g["A"]) < 1e-13
    R = OF.compute_RHS(fm, x, L, dt, c0, c1, force=g["force"], flow=g["flow"])
    assert rel2(R, g["RHS"]) < 1e-13
    A_sq, b = OF.apply_BC_rectangular(fm, x, L, E, A, R, end_force=g["force"][-1])
    assert rel2(A_sq, g["A_sq"]) < 1e-13
    assert rel2(b, g["b"]) < 1e-13
    assert rel2(OF.force_operator(fm, x, L, E), g["F"]) < 1e-13
    # minus end clamped to a body: 'velocity' + 'angular_velocity' rows
    A_cl, b_cl = OF.apply_BC_rectangular(fm, x, L, E, A, R, end_force=g["force"][-1],
                                         clamped=True, dt=dt, c0=c0, flow=g["flow"],
                                         force=g["force"])
    assert rel2(A_cl, g["A_sq_clamped"]) < 1e-13
    assert rel2(b_cl, g["b_clamped"]) < 1e-13
    # only the seven minus-end rows differ from the free fiber
    assert np.array_equal(A_cl[:4 * N - 14], A_sq[:4 * N - 14])
    assert np.array_equal(A_cl[4 * N - 7:], A_sq[4 * N - 7:])
    # plus end hinged at the periphery: 'velocity' + 'torque' rows at s = L,
    # no prescribed end force
    A_hg, b_hg = OF.apply_BC_rectangular(fm, x, L, E, A, R, hinged=True, dt=dt, c0=c0,
                                         flow=g["flow"], force=g["force"])
    assert rel2(A_hg, g["A_sq_hinged"]) < 1e-13
    assert rel2(b_hg, g["b_hinged"]) < 1e-13


@pytest.mark.parametrize("N", [16, 32, 64])
def test_fd_weights_polynomial_exact(N):
    fm = OF.FiberMatrices(N)
    a = fm.alpha
    for k, D in enumerate(fm.D0, start=1):
        np.testing.assert_allclose(D @ np.ones(N), 0.0, atol=1e-8 * np.abs(D).max())
        # stencil of nfd + k = 4 + k points is exact for degree < 4 + k
        for p in range(k, 4 + k):
            exact = np.prod(np.arange(p, p - k, -1)) * a ** (p - k)
            assert rel2(D @ a**p, exact) < 1e-9


def test_fd_analytic_derivatives_like_reference_test():
    # f = cos(pi s) + s + exp(2 s) on s in [-1, 1] (tests/test_finite_diff.py:50-54)
    N = 96
    fm = OF.FiberMatrices(N)
    s = fm.alpha
    f = np.cos(np.pi * s) + s + np.exp(2 * s)
    exact = [-np.pi * np.sin(np.pi * s) + 1 + 2 * np.exp(2 * s),
             -np.pi**2 * np.cos(np.pi * s) + 4 * np.exp(2 * s),
             np.pi**3 * np.sin(np.pi * s) + 8 * np.exp(2 * s),
             np.pi**4 * np.cos(np.pi * s) + 16 * np.exp(2 * s)]
    for k, (D, e) in enumerate(zip(fm.D0, exact)):
        assert rel2(D @ f, e) < [1e-5, 1e-4, 1e-3, 1e-2][k]


def test_barycentric_properties():
    x = np.linspace(-1, 1, 20)
    y = np.linspace(-0.97, 0.97, 33)
    P = OF.barycentric_matrix(x, y)
    np.testing.assert_allclose(P @ np.ones(20), 1.0, rtol=1e-14)
    np.testing.assert_allclose(OF.barycentric_matrix(x, x), np.eye(20), atol=0)
    # Gaussian resampled on finer grid (tests/test_barycentric.py:19-20)
    xg = np.linspace(-1, 1, 64)
    yg = np.linspace(-1, 1, 101)
    g = np.exp(-4 * xg**2)
    assert rel2(OF.barycentric_matrix(xg, yg) @ g, np.exp(-4 * yg**2)) < 1e-3


def test_chebyshev_variant_spectral():
    fm = OF.FiberMatrices(24, kind="chebyshev")
    a = fm.alpha
    f = np.sin(2 * a)
    assert rel2(fm.D0[0] @ f, 2 * np.cos(2 * a)) < 1e-10
    assert abs(fm.weights0.sum() - 2.0) < 1e-13
    # first-kind collocation grids interpolate a polynomial exactly
    assert rel2(fm.P_X @ a**5, fm.alpha_roots**5) < 1e-12


def test_stencil_width_out_of_range_is_rejected():
    """num_points_finite_diff outside [1, N - 4] leaves the k-th derivative
    without its nfd + k nodes: rejected with a clear error (the reference's
    Fornberg loop fails with an IndexError there)."""
    import pytest
    from fiber3d_amd.fiber_matrices import FiberMatrices
    FiberMatrices(16, 12)
    for bad in (0, 13):
        with pytest.raises(ValueError):
            FiberMatrices(16, bad)
