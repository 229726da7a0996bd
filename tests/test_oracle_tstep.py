"""Pins of the system-level oracle (matvec, preconditioner, GMRES)."""
import numpy as np
import pytest

from conftest import rel2
from fiber3d_amd import configs
from oracle import tstep as OT


@pytest.fixture(scope="module")
def small_system():
    spec = configs.FiberSpec(6, 16, 1.0, 2.0)
    fibers = configs.make_fibers(spec, seed=7)
    shell = configs.make_periphery(240, 3.5)
    return OT.HydroSystem(fibers, shell, eta=1.0, dt=1e-3, E=0.01)


def test_matvec_equals_dense_operator(small_system):
    S = small_system
    A = S.dense_operator()
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.normal(size=S.size)
        assert rel2(S.matvec(x), A @ x) < 1e-12


def test_preconditioner_inverts_its_blocks(small_system):
    S = small_system
    rng = np.random.default_rng(1)
    x = rng.normal(size=S.size)
    # P applied blockwise, then P^-1: identity
    Px = np.empty_like(x)
    Px[:3 * S.nper] = S.M_shell @ x[:3 * S.nper]
    for k in range(len(S.fibers)):
        N = S.sizes[k]
        o = 3 * S.nper + 4 * S.offs[k]
        Px[o:o + 4 * N] = S.A_sq[k] @ x[o:o + 4 * N]
    assert rel2(S.precond(Px), x) < 1e-10


def test_gmres_matches_direct_solve(small_system):
    S = small_system
    A = S.dense_operator()
    xd = np.linalg.solve(A, S.rhs)
    x, info, it = OT.gmres(S.matvec, S.rhs, S.precond, tol=1e-10)
    assert info == 0 and it > 0
    assert np.linalg.norm(S.rhs - A @ x) <= 1e-10 * np.linalg.norm(S.rhs) * 1.0001
    assert rel2(x, xd) < 1e-7


def test_single_fiber_no_shell_reduces_to_block():
    # one fiber and no shell: the matvec is A_sq XT minus the (empty) flow term
    spec = configs.FiberSpec(1, 24, 1.0, 0.5)
    S = OT.HydroSystem(configs.make_fibers(spec, seed=3))
    x = np.random.default_rng(2).normal(size=S.size)
    assert rel2(S.matvec(x), S.A_sq[0] @ x) < 1e-14


def test_lone_sphere_obeys_stokes_law():
    """A free sphere under an external force F and torque L translates at
    F/(6 pi eta a) and rotates at L/(8 pi eta a^3): pins the body rows
    (double layer, singularity subtraction, K) and the Stokeslet + rotlet
    completion at the centre."""
    a, eta = 0.5, 1.3
    body = configs.make_sphere_body([0.3, -0.2, 0.1], a, 800)
    F = np.array([0.7, -0.4, 1.1])
    Lt = np.array([-0.5, 0.9, 0.2])
    S = OT.HydroSystem([], None, eta=eta, bodies=[body], body_force=[np.concatenate([F, Lt])])
    x, info, it = OT.gmres(S.matvec, S.rhs, S.precond, tol=1e-10)
    assert info == 0
    _, vel = S.body_parts(x)
    U, W = vel[0][:3], vel[0][3:]
    assert rel2(U, F / (6 * np.pi * eta * a)) < 1e-2
    assert rel2(W, Lt / (8 * np.pi * eta * a**3)) < 1e-2


def test_body_preconditioner_inverts_its_block():
    body = configs.make_sphere_body([0.0, 0.0, 0.0], 0.5, 200)
    S = OT.HydroSystem([], None, bodies=[body])
    x = np.random.default_rng(4).normal(size=S.size)
    # the body block alone is the whole operator here
    assert rel2(S.precond(S.matvec(x)), x) < 1e-9
