"""Fibers clamped to rigid bodies (tstep.py:878-881, tstep_utils.build_link_matrix
:948-1091) on the GPU against the oracle: right-hand side, operator in FP64
and FP32-compensated, preconditioner, and time steps that keep every minus
end on its (moving) nucleation site."""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd.gmres import gmres
from fiber3d_amd.learned import LearnedGuess
from fiber3d_amd.system import FiberSystem, Shell
from fiber3d_amd.tstep import TimeStepper
from oracle import tstep as OT

pytestmark = pytest.mark.gpu


def _setup():
    a = 0.5
    centres = [np.array([-1.0, 0.2, 0.0]), np.array([1.2, -0.3, 0.4])]
    bodies = [configs.make_sphere_body(c, a, 150) for c in centres]
    rng = np.random.default_rng(3)
    fibers, attached = [], []
    for j, c in enumerate(centres):
        for _ in range(2):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            v = np.cross(u, rng.normal(size=3))
            v /= np.linalg.norm(v)
            N = 16 if j == 0 else 24
            s = np.linspace(0.0, 0.8, N)
            # slightly curved, so the x_ss terms of the clamped rows are live
            fibers.append(c + (a + s)[:, None] * u + (0.05 * s * s)[:, None] * v)
            attached.append((len(fibers) - 1, j, a * u))
    # one free fiber between the bodies
    s = np.linspace(0.0, 0.8, 16)
    fibers.append(np.stack([0.1 + 0 * s, -1.2 + s, 0.9 + 0 * s], axis=1))
    shell = configs.make_periphery(300, 3.2)
    return fibers, bodies, attached, shell


def test_attached_system_matches_oracle():
    print("case attached vs oracle", flush=True)
    fibers, bodies, attached, (xs, ns, ws) = _setup()
    force = [np.array([0.3, -0.2, 0.1, 0.05, 0.0, -0.1]), np.zeros(6)]
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), bodies=bodies, body_force=force,
                      attached=attached)
    ora = OT.HydroSystem(fibers, (xs, ns, ws), bodies=bodies, body_force=force,
                         attached=attached)
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-11
    x = np.random.default_rng(4).normal(size=ora.size)
    xt = torch.from_numpy(x).cuda()
    ref = ora.matvec(x)
    assert rel2(gpu.matvec(xt).cpu().numpy(), ref) < 1e-10
    assert rel2(gpu.matvec(xt, precision="f32c").cpu().numpy(), ref) < 1e-6
    # the clamped rows enter the fiber blocks and therefore the preconditioner
    assert rel2(gpu.precond(xt).cpu().numpy(), ora.precond(x)) < 1e-9
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-10)
    torch.cuda.synchronize()
    assert res.info == 0
    xo = res.x.cpu().numpy()
    assert np.linalg.norm(ora.rhs - ora.matvec(xo)) / np.linalg.norm(ora.rhs) < 1.05e-10


def test_attached_time_steps_keep_links():
    print("case attached time steps", flush=True)
    fibers, bodies, attached, (xs, ns, ws) = _setup()
    force = [np.array([0.5, 0.0, 0.0, 0.0, 0.0, 0.3]), np.array([0.0, -0.4, 0.0, 0.2, 0.0, 0.0])]
    ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-8, bodies=bodies,
                     body_force=force, attached=attached, record=True,
                     predictor=LearnedGuess(use_net=False))
    S = ts.sys
    for _ in range(6):
        st = ts.step()
        assert st.residual <= 1e-8
        for k, j, link in attached:
            site = S.bodies.centres[j] + S.bodies.rot[j] @ torch.from_numpy(link).cuda()
            assert torch.linalg.norm(S.X[S.offs[k]] - site) < 1e-12
    torch.cuda.synchronize()
    true, drhs = ts.verify()
    assert max(true) <= 1e-8 and drhs == 0.0
    # the bodies moved and turned, and the attached fibers with them
    assert float(torch.linalg.norm(S.bodies.centres[0] - torch.from_numpy(bodies[0][0]).cuda())) > 0
    assert float(torch.linalg.norm(S.bodies.rot[0] - torch.eye(3, device="cuda",
                                                               dtype=torch.float64))) > 0
