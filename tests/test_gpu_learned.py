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


def test_learned_guess_beats_extrapolation_on_held_out_steps():
    """A time series: the learned predictor trains its fiber net on its
    first `train` steps (picking its baseline on the last 4 of them) and
    checks it on 4 gate steps; then all predictors run the same number of
    further (held-out) steps.  Here (dense shell preconditioner) the baseline is the
    extrapolation of the residual-corrected history (x + P^-1 r); on the
    held-out steps the learned guesses have at least 10% smaller residuals
    and need at least 10% fewer GMRES iterations than the plain
    extrapolation, and never do worse than the corrected extrapolation alone:
    the net is kept only if it beats its baseline on both gates (validation
    MSE, then guess residuals, below 0.9 times the baseline's)."""
    print("case learned time series", flush=True)
    fibers, shell = _setup(nfib=160, nper=1600, seed=3)
    train, held = 30, 40
    runs, r0s, reports = {}, {}, {}
    preds = (("zero", None),
             ("extrapolated", LearnedGuess(train_steps=10**9, use_net=False)),
             ("corrected", LearnedGuess(train_steps=10**9, use_net=False, correct=True)),
             ("learned", LearnedGuess(train_steps=train, epochs=60)))
    for name, pred in preds:
        ts = TimeStepper(fibers, shell, mode="fast", tol=1e-8, predictor=pred)
        warm = preds[-1][1].warm_steps  # the learned predictor's whole warm-up window
        stats = [ts.step() for _ in range(warm + held)]
        assert all(st.residual <= 1e-8 for st in stats)
        tail = stats[warm:]
        runs[name] = float(np.mean([st.iterations for st in tail]))
        r0s[name] = float(np.mean([st.guess_residual for st in tail])) if pred else None
        if pred is not None and pred.use_net:
            assert pred.trained
            reports[name] = pred.report
    print("held-out mean GMRES iterations:", runs, "guess residuals:", r0s, reports, flush=True)
    assert runs["extrapolated"] < runs["zero"]
    assert r0s["learned"] <= 0.9 * r0s["extrapolated"]
    assert runs["learned"] <= 0.9 * runs["extrapolated"]
    # the gated net never makes the corrected extrapolation worse (the bar
    # leaves room for the trajectories' atomic-summation differences)
    assert r0s["learned"] <= 1.05 * r0s["corrected"]
    assert reports["learned"]["baseline"]["mode"] == "corrected"
    rep = reports["learned"]["fiber"]
    if rep["enabled"]:
        assert rep["val_mse_net"] < 0.9 * rep["val_mse_baseline"]
        assert rep["held_out_guess_residual_net"] < 0.9 * rep["held_out_guess_residual_baseline"]
    elif "held_out_guess_residual_net" in rep:
        assert rep["held_out_guess_residual_net"] >= 0.9 * rep["held_out_guess_residual_baseline"]
    else:
        assert rep["val_mse_net"] >= 0.9 * rep["val_mse_baseline"]
