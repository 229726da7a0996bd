"""Fibers whose plus end is hinged, as a fiber that reached the periphery
(tstep.py:886-888; fiber.py:940-960 for the rows, tstep.py:1167-1168 for the
flow term), and fibers clamped to the cortex (tstep.py:881-883, flow term
:1162-1166), on the GPU against the oracle: right-hand side, operator in FP64
and FP32-compensated, preconditioner, and time steps that hold every hinged
or clamped end in place."""
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

R_SHELL = 2.0


def _setup():
    """Two spheres, one fiber clamped to the first with its plus end hinged
    on the periphery, two free fibers reaching the periphery (hinged) and
    one free fiber inside (not hinged)."""
    a = 0.4
    centres = [np.array([-0.6, 0.1, 0.0]), np.array([0.5, -0.5, 0.3])]
    bodies = [configs.make_sphere_body(c, a, 150) for c in centres]
    rng = np.random.default_rng(11)
    fibers, attached, hinged = [], [], []

    def towards_shell(p0, N, bend):
        u = p0 / np.linalg.norm(p0)
        L = R_SHELL - 0.05 - np.linalg.norm(p0)
        v = np.cross(u, rng.normal(size=3))
        v /= np.linalg.norm(v)
        s = np.linspace(0.0, 1.0, N)
        return p0 + (L * s)[:, None] * u + (bend * s * (1 - s))[:, None] * v

    u = np.array([-1.0, 0.2, 0.1])
    u /= np.linalg.norm(u)
    fibers.append(towards_shell(centres[0] + a * u, 16, 0.05))
    attached.append((0, 0, a * u))
    hinged.append(0)
    for p0, N in [(np.array([0.2, 0.9, -0.3]), 24), (np.array([0.1, -0.2, -1.0]), 16)]:
        fibers.append(towards_shell(p0, N, 0.08))
        hinged.append(len(fibers) - 1)
    s = np.linspace(0.0, 0.6, 16)
    fibers.append(np.stack([0.6 + 0 * s, 0.4 + s, -0.2 + 0 * s], axis=1))
    shell = configs.make_periphery(400, R_SHELL)
    lengths = [float(np.sum(np.linalg.norm(np.diff(x, axis=0), axis=1))) for x in fibers]
    return fibers, bodies, attached, hinged, shell, lengths


@pytest.mark.parametrize("cortex", [False, True])
def test_hinged_system_matches_oracle(cortex):
    print(f"case hinged vs oracle cortex={cortex}", flush=True)
    fibers, bodies, attached, hinged, (xs, ns, ws), lengths = _setup()
    force = [np.array([0.3, -0.2, 0.1, 0.05, 0.0, -0.1]), np.array([0.0, 0.2, 0.0, 0.0, 0.1, 0.0])]
    kw = dict(bodies=bodies, body_force=force, attached=attached, hinged=hinged, lengths=lengths,
              cortex=cortex)
    gpu = FiberSystem(fibers, Shell(xs, ns, ws), **kw)
    ora = OT.HydroSystem(fibers, (xs, ns, ws), **kw)
    assert rel2(gpu.rhs.cpu().numpy(), ora.rhs) < 1e-11
    x = np.random.default_rng(4).normal(size=ora.size)
    xt = torch.from_numpy(x).cuda()
    ref = ora.matvec(x)
    assert rel2(gpu.matvec(xt).cpu().numpy(), ref) < 1e-10
    assert rel2(gpu.matvec(xt, precision="f32c").cpu().numpy(), ref) < 1e-6
    # the hinged rows enter the fiber blocks and therefore the preconditioner
    assert rel2(gpu.precond(xt).cpu().numpy(), ora.precond(x)) < 1e-9
    res = gmres(gpu.matvec, gpu.rhs, gpu.precond, tol=1e-10)
    torch.cuda.synchronize()
    assert res.info == 0
    xo = res.x.cpu().numpy()
    assert np.linalg.norm(ora.rhs - ora.matvec(xo)) / np.linalg.norm(ora.rhs) < 1.05e-10


def test_hinged_time_steps_hold_the_plus_ends():
    print("case hinged time steps", flush=True)
    fibers, bodies, attached, hinged, (xs, ns, ws), lengths = _setup()
    force = [np.array([0.5, 0.0, 0.0, 0.0, 0.0, 0.3]), np.array([0.0, -0.4, 0.0, 0.2, 0.0, 0.0])]
    ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-8, bodies=bodies,
                     body_force=force, attached=attached, hinged=hinged, lengths=lengths,
                     record=True, predictor=LearnedGuess(use_net=False))
    S = ts.sys
    ends = {k: S.X[S.offs[k + 1] - 1].clone() for k in hinged}
    free_end = S.X[S.offs[4] - 1].clone()
    for _ in range(6):
        st = ts.step()
        assert st.residual <= 1e-8
        for k in hinged:
            assert torch.linalg.norm(S.X[S.offs[k + 1] - 1] - ends[k]) < 1e-6
    torch.cuda.synchronize()
    true, drhs = ts.verify()
    assert max(true) <= 1e-8 and drhs == 0.0
    # a fiber held at both ends near the periphery: the FP32-compensated
    # operator's error on the corrections is measured above the 1e-6 floor,
    # and every step's certificate bounds its true residual
    print("inner_rel_err", ts.inner_rel_err, "kinds", [h.residual_kind for h in ts.history])
    assert ts.inner_rel_err > 1e-6
    for h, t in zip(ts.history, true):
        assert t <= h.residual_bound * (1 + 1e-9) + 1e-15
    # the unhinged fiber's plus end is carried by the flow
    assert float(torch.linalg.norm(S.X[S.offs[4] - 1] - free_end)) > 1e-4


def test_cortex_time_steps_hold_the_minus_ends():
    print("case cortex time steps", flush=True)
    fibers, bodies, attached, hinged, (xs, ns, ws), lengths = _setup()
    force = [np.array([0.5, 0.0, 0.0, 0.0, 0.0, 0.3]), np.array([0.0, -0.4, 0.0, 0.2, 0.0, 0.0])]
    ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-8, bodies=bodies,
                     body_force=force, attached=attached, lengths=lengths, cortex=True,
                     record=True, predictor=LearnedGuess(use_net=False))
    S = ts.sys
    cortex = [1, 2, 3]
    ends = {k: S.X[S.offs[k]].clone() for k in cortex}
    tang = {k: S.xs[S.offs[k]].clone() for k in cortex}
    tip = S.X[S.offs[4] - 1].clone()
    for _ in range(6):
        st = ts.step()
        assert st.residual <= 1e-8
        for k in cortex:
            assert torch.linalg.norm(S.X[S.offs[k]] - ends[k]) < 1e-6
            assert torch.linalg.norm(S.xs[S.offs[k]] - tang[k]) < 1e-5
    torch.cuda.synchronize()
    true, drhs = ts.verify()
    assert max(true) <= 1e-8 and drhs == 0.0
    for h, t in zip(ts.history, true):
        assert t <= h.residual_bound * (1 + 1e-9) + 1e-15
    # the free plus end of a clamped fiber still moves with the flow
    assert float(torch.linalg.norm(S.X[S.offs[4] - 1] - tip)) > 1e-4


def test_hinged_adaptive_rejections_keep_the_ends():
    """Adaptive steps with a tolerance tight enough to force rejections: a
    rejected step restores the configuration and re-assembles the hinged and
    clamped rows at dt / 2, and the hinged ends stay put throughout."""
    print("case hinged adaptive", flush=True)
    fibers, bodies, attached, hinged, (xs, ns, ws), lengths = _setup()
    force = [np.array([0.5, 0.0, 0.0, 0.0, 0.0, 0.3]), np.array([0.0, -0.4, 0.0, 0.2, 0.0, 0.0])]
    ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-8, bodies=bodies,
                     body_force=force, attached=attached, hinged=hinged, lengths=lengths,
                     cortex=True, adaptive=True, dt=4e-3, tol_tstep=1e-9, dt_min=1e-3,
                     periphery_radius=R_SHELL, record=True,
                     predictor=LearnedGuess(use_net=False))
    S = ts.sys
    ends = {k: S.X[S.offs[k + 1] - 1].clone() for k in hinged}
    for _ in range(5):
        st = ts.step()
        assert st.residual <= 1e-8
        for k in hinged:
            assert torch.linalg.norm(S.X[S.offs[k + 1] - 1] - ends[k]) < 1e-6
    dts = [h.dt for h in ts.history]
    print("dt per solve", dts, "rejected", ts.rejected, flush=True)
    assert ts.rejected >= 1 and min(dts) < dts[0]
    torch.cuda.synchronize()
    true, drhs = ts.verify()
    assert max(true) <= 1e-8 and drhs == 0.0
