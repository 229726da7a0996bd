"""Adaptive time stepping (tstep.py:427-581, check_error :1874-1925): the
inextensibility error against the oracle's derivative matrices, rejected
steps restore the configuration and halve dt, accepted ones grow it."""
import numpy as np
import pytest
import torch

from fiber3d_amd import configs
from fiber3d_amd.learned import LearnedGuess
from fiber3d_amd.system import Shell
from fiber3d_amd.tstep import TimeStepper
from oracle import fiber as OF

pytestmark = pytest.mark.gpu


def _setup():
    fibers = configs.make_fibers(configs.FiberSpec(6, 16, 1.0, 2.0), seed=9)
    xs, ns, ws = configs.make_periphery(300, 4.0)
    return fibers, Shell(xs, ns, ws)


def test_inextensibility_error_matches_oracle():
    print("case inextensibility error", flush=True)
    fibers, shell = _setup()
    ts = TimeStepper(fibers, shell, mode="plain", tol=1e-10)
    ts.step()
    torch.cuda.synchronize()
    X, _ = ts.fiber_arrays()
    fm = OF.FiberMatrices(16)
    ref = max(np.abs(np.linalg.norm(fm.Ds(1.0)[0] @ x, axis=1) - 1.0).max() for x in X)
    assert ref > 0.0
    assert abs(ts.inextensibility_error() - ref) <= 1e-12 + 1e-9 * ref


def test_rejected_steps_restore_and_halve_dt():
    print("case adaptive reject", flush=True)
    fibers, shell = _setup()
    ts = TimeStepper(fibers, shell, mode="fast", dt=1e-3, tol=1e-8, adaptive=True,
                     tol_tstep=1e-14, dt_min=2.5e-4, predictor=LearnedGuess(use_net=False))
    X0 = ts.sys.X.clone()
    st = ts.step()
    assert not st.accepted and st.dt == 1e-3 and ts.sys.dt == 5e-4
    assert torch.equal(ts.sys.X, X0) and ts.time == 0.0
    st = ts.step()
    assert not st.accepted and ts.sys.dt == 2.5e-4 and torch.equal(ts.sys.X, X0)
    st = ts.step()      # at dt_min: accepted whatever the error
    assert st.accepted and st.dt == 2.5e-4 and ts.time == 2.5e-4
    assert not torch.equal(ts.sys.X, X0) and ts.rejected == 2
    assert all(s.residual <= 1e-8 for s in ts.history)


def test_accepted_steps_grow_dt_and_track_fixed_steps():
    print("case adaptive grow", flush=True)
    fibers, shell = _setup()
    ts = TimeStepper(fibers, shell, mode="fast", dt=1e-3, tol=1e-8, adaptive=True,
                     tol_tstep=1.0, dt_max=1.5e-3, predictor=LearnedGuess(use_net=False),
                     periphery_radius=4.0)
    dts = [ts.step().dt for _ in range(8)]
    torch.cuda.synchronize()
    assert np.allclose(dts[:3], [1e-3, 1.2e-3, 1.44e-3]) and np.allclose(dts[3:], 1.5e-3)
    assert all(s.accepted and s.residual <= 1e-8 for s in ts.history)
    assert abs(ts.time - sum(dts)) < 1e-15
    # the time-aware extrapolation keeps the guess useful across the changing steps
    assert all(s.guess_used for s in ts.history[2:])
    assert np.mean([s.iterations for s in ts.history[-4:]]) < ts.history[0].iterations
