"""A time step of the reference's example structure read from its own file,
written back in the reference's output formats."""
import io
import os

import numpy as np
import pytest
import torch

from fiber3d_amd import configs
from fiber3d_amd.fileio import fibers_from_file, read_fibers_file
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
