"""The fast mode's closing residual: the FP64-anchored update (one FP64
product per solve) against the recomputed FP64 residual, and every recorded
step re-assembled and re-checked in FP64 (TimeStepper.verify)."""
import pytest
import torch

from conftest import rel2
from fiber3d_amd import configs
from fiber3d_amd.learned import LearnedGuess
from fiber3d_amd.system import Shell
from fiber3d_amd.tstep import TimeStepper

pytestmark = pytest.mark.gpu


def _run(certify, steps=6):
    fibers = configs.make_fibers(configs.FiberSpec(6, 16, 1.0, 2.0), seed=9)
    bodies = configs.make_bodies(2, 0.5, 120, 2.2, fibers=fibers, seed=1)
    xs, ns, ws = configs.make_periphery(300, 4.0)
    ts = TimeStepper(fibers, Shell(xs, ns, ws), mode="fast", tol=1e-8, bodies=bodies,
                     predictor=LearnedGuess(train_steps=10**9, use_net=False),
                     certify=certify, record=True)
    stats = [ts.step() for _ in range(steps)]
    torch.cuda.synchronize()
    return ts, stats


def test_update_residual_matches_fp64_certificate():
    print("case certify update vs fp64", flush=True)
    ts_u, st_u = _run("update")
    ts_f, st_f = _run("fp64")
    # the closing residual does not change the iterate: the same trajectory
    # (up to the rounding of separately built preconditioners, which moves
    # where GMRES stops inside the 1e-8 tolerance: 5e-12 measured)
    assert [s.iterations for s in st_u] == [s.iterations for s in st_f]
    assert rel2(ts_u.sys.X.cpu().numpy(), ts_f.sys.X.cpu().numpy()) < 1e-10
    assert rel2(ts_u.sys.bodies.pos.cpu().numpy(), ts_f.sys.bodies.pos.cpu().numpy()) < 1e-10
    kinds = [s.residual_kind for s in st_u]
    # FP64 until the operator's error has been measured on a correction
    assert kinds[:2] == ["fp64", "fp64"] and "update" in kinds, kinds
    assert ts_u.inner_rel_err is not None
    assert all(s.residual_kind == "fp64" for s in st_f)
    true_u, drhs_u = ts_u.verify()
    true_f, drhs_f = ts_f.verify()
    # re-assembly at a recorded geometry reproduces its right-hand side exactly
    assert drhs_u == 0.0 and drhs_f == 0.0
    for s, t in zip(st_u, true_u):
        assert t <= 1e-8
        assert t <= s.residual_bound * (1 + 1e-9) + 1e-15
        if s.residual_kind == "update":
            assert abs(t - s.residual) <= 1e-11, (t, s.residual)
    for s, t in zip(st_f, true_f):
        assert abs(t - s.residual) <= 1e-14 + 1e-6 * t


def test_verify_restores_the_current_state():
    print("case verify restores state", flush=True)
    ts, _ = _run("update", steps=3)
    S = ts.sys
    X, pos, rhs = S.X.clone(), S.bodies.pos.clone(), S.rhs.clone()
    ts.verify()
    assert torch.equal(S.X, X) and torch.equal(S.bodies.pos, pos)
    assert torch.equal(S.rhs, rhs)
    st = ts.step()
    assert st.residual <= 1e-8
