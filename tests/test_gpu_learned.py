"""Learned GMRES initial guess: fallback safety and gain on a short time series."""
import numpy as np
import pytest
import torch

from fiber3d_amd import configs
from fiber3d_amd.gmres import gmres
from fiber3d_amd.learned import LearnedGuess
from fiber3d_amd.system import FiberSystem, Shell
from fiber3d_amd.tstep import TimeStepper

pytestmark = pytest.mark.gpu


def _setup(nfib=48, nper=600, seed=3):
    spec = configs.FiberSpec(nfib, 32, 1.0, 3.0)
    fibers = configs.make_fibers(spec, seed)
    xs, ns, ws = configs.make_periphery(nper, 4.5)
    return fibers, Shell(xs, ns, ws)


def test_bad_guess_falls_back_to_zero_start():
    print("case learned fallback", flush=True)
    fibers, shell = _setup(16, 300)
    S = FiberSystem(fibers, shell)
    plain = gmres(S.matvec, S.rhs, S.precond, tol=1e-8)
    nan_guess = torch.full_like(S.rhs, float("nan"))
    r1 = gmres(S.matvec, S.rhs, S.precond, tol=1e-8, x0=nan_guess)
    huge = torch.full_like(S.rhs, 1e12)
    r2 = gmres(S.matvec, S.rhs, S.precond, tol=1e-8, x0=huge)
    torch.cuda.synchronize()
    for r in (r1, r2):
        assert not r.guess_used and r.info == 0 and r.residual <= 1e-8
        assert r.iterations == plain.iterations
    # a good guess (the solution itself) converges at once
    r3 = gmres(S.matvec, S.rhs, S.precond, tol=1e-8, x0=plain.x)
    assert r3.guess_used and r3.iterations == 0


def test_learned_guess_reduces_iterations():
    print("case learned time series", flush=True)
    fibers, shell = _setup()
    train = 24
    runs, reports = {}, {}
    for name, pred in (("zero", None),
                       ("extrapolated", LearnedGuess(train_steps=10**9, use_net=False)),
                       ("learned", LearnedGuess(train_steps=train, epochs=60))):
        ts = TimeStepper(fibers, shell, mode="fast", tol=1e-8, predictor=pred)
        its = []
        for _ in range(train + 16):
            st = ts.step()
            assert st.residual <= 1e-8
            its.append(st.iterations)
        runs[name] = float(np.mean(its[-8:]))
        if pred is not None and pred.use_net:
            assert pred.trained
            reports[name] = pred.report
    print("mean GMRES iterations over the last 8 steps:", runs, reports)
    assert runs["extrapolated"] < runs["zero"]
    # the gate keeps a net only if it beats the extrapolation on held-out steps
    for part in reports["learned"].values():
        if part["enabled"]:
            vals = list(part.values())
            assert vals[0] < 0.9 * vals[1]
    assert runs["learned"] <= runs["extrapolated"] + 0.25
