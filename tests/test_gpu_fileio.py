"""A time step of the reference's example structure read from its own file,
written back in the reference's output formats."""
import io
import os

import numpy as np
import pytest
import torch

from fiber3d_amd import configs
from fiber3d_amd.fileio import fibers_from_file, read_clones_file, read_fibers_file
from fiber3d_amd.system import Shell
from fiber3d_amd.tstep import TimeStepper

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_step_reference_structure_and_write_outputs(tmp_path):
    print("case fileio two.fibers", flush=True)
    xs, Es, Ls = fibers_from_file(os.path.join(GOLDEN, "two.fibers"))
    # the structure spans z in [0.15, 5.85]: a periphery of radius 8 holds it
    xsh, nsh, wsh = configs.make_periphery(800, 8.0)
    # the example's nodes are Chebyshev-Lobatto points along each fiber
    ts = TimeStepper(xs, Shell(xsh, nsh, wsh), mode="plain", E=Es[0], lengths=Ls, tol=1e-10,
                     discretisation="chebyshev")
    st = ts.step()
    torch.cuda.synchronize()
    assert st.residual <= 1e-10
    buf = io.BytesIO()
    ts.write_frame(buf)
    rows = np.loadtxt(io.StringIO(buf.getvalue().decode()))
    X, T = ts.fiber_arrays()
    assert np.array_equal(rows[1], [32, Es[0], Ls[0], 0])
    assert np.array_equal(rows[2:34, :3], X[0]) and np.array_equal(rows[2:34, 3], T[0])
    name = tmp_path / "out_fibers_resume.fibers"
    ts.write_resume(name)
    info, coor = read_fibers_file(name)
    assert [h[:3] for h in info] == [[32, Es[0], Ls[0]], [32, Es[1], Ls[1]]]
    assert np.array_equal(coor, np.concatenate(X))
    # the step moved the fibers, and by little
    moved = np.abs(coor - np.concatenate(xs)).max()
    assert 0.0 < moved < 1e-2


def _quat_matrix(q):
    """Rotation matrix of a unit quaternion (s, p) (quaternion.py:41-50)."""
    s, p = q[0], q[1:]
    return 2.0 * (np.outer(p, p) + (s * s - 0.5) * np.eye(3)
                  + s * np.array([[0.0, -p[2], p[1]], [p[2], 0.0, -p[0]], [-p[1], p[0], 0.0]]))


def test_body_orientations_track_rotation_and_write(tmp_path):
    """The bodies' quaternions advance like the reference's (q <- q(Omega dt) q)
    and stay equal to the tracked rotation times the initial orientation."""
    print("case body quaternions", flush=True)
    fibers = configs.make_fibers(configs.FiberSpec(4, 16, 1.0, 2.0), seed=9)
    bodies = configs.make_bodies(2, 0.5, 120, 2.2, fibers=fibers, seed=1)
    q0 = np.array([[0.9, 0.1, -0.3, 0.2], [1.0, 0.0, 0.0, 0.0]])
    q0 /= np.linalg.norm(q0, axis=1, keepdims=True)
    xsh, nsh, wsh = configs.make_periphery(300, 4.0)
    force = [np.array([0.0, 0.0, 0.0, 1.0, -0.5, 0.2]), np.array([0.3, 0.0, 0.0, 0.0, 0.8, 0.0])]
    ts = TimeStepper(fibers, Shell(xsh, nsh, wsh), mode="plain", bodies=bodies,
                     body_force=force, body_orientations=q0, dt=0.01)
    for _ in range(4):
        ts.step()
    torch.cuda.synchronize()
    B = ts.sys.bodies
    q = B.quat.cpu().numpy()
    R = B.rot.cpu().numpy()
    for k in range(2):
        assert abs(np.linalg.norm(q[k]) - 1.0) < 1e-14
        assert np.abs(_quat_matrix(q[k]) - R[k] @ _quat_matrix(q0[k])).max() < 1e-13
    assert np.abs(q - q0).max() > 1e-4          # they did turn
    name = tmp_path / "out_bodies_resume.clones"
    ts.write_resume_bodies(name)
    n, loc, qq = read_clones_file(name)
    assert n == 2 and np.array_equal(loc, B.centres.cpu().numpy())
    assert np.abs(qq - q).max() < 1e-15
    buf = io.BytesIO()
    vbuf = io.BytesIO()
    ts.write_bodies(buf, vbuf)
    rows = np.loadtxt(io.StringIO(buf.getvalue().decode()))
    assert rows.shape == (5, 7) and abs(rows[1, 0] - 0.5) < 1e-12
    vel = np.loadtxt(io.StringIO(vbuf.getvalue().decode()))
    assert np.array_equal(vel[1:], ts.body_velocity.cpu().numpy())


def test_stepper_from_input_file(tmp_path):
    """stepper_from_input: an input file in the reference's format naming
    the reference's example fibers sets up the stepper (dt, eta, tolerance,
    fibers, E and lengths from the headers), which then steps to its
    tolerance."""
    print("case stepper from input", flush=True)
    from fiber3d_amd.fileio import ReadInput
    from fiber3d_amd.tstep import stepper_from_input
    f = tmp_path / "data.main"
    f.write_text("dt 1e-3\neta 1.0\nsolver_tolerance 1e-10\nE 0.5\n"
                 "num_points_finite_diff 4\n"
                 "structure_fibers %s\n" % os.path.join(GOLDEN, "two.fibers"))
    r = ReadInput(str(f))
    xsh, nsh, wsh = configs.make_periphery(800, 8.0)
    ts = stepper_from_input(r, Shell(xsh, nsh, wsh), mode="plain", discretisation="chebyshev")
    xs, Es, Ls = fibers_from_file(os.path.join(GOLDEN, "two.fibers"), E=0.5)
    assert ts.sys.dt == 1e-3 and ts.tol == 1e-10 and ts.sys.E == Es[0]
    np.testing.assert_array_equal(ts.sys.lengths_h, Ls)
    np.testing.assert_array_equal(ts.sys.X.cpu().numpy(), np.concatenate(xs))
    for _ in range(2):
        assert ts.step().residual <= 1e-10
