"""Host checks of the initial-guess extrapolation weights."""
import numpy as np
import pytest
import torch

from fiber3d_amd.learned import lsq_weights


@pytest.mark.parametrize("npts,deg", [(2, 1), (3, 2), (5, 2), (6, 2), (8, 3)])
def test_lsq_weights_reproduce_polynomials(npts, deg):
    """The one-step-ahead value of any polynomial of degree <= deg is exact."""
    w = lsq_weights(npts, deg)
    rng = np.random.default_rng(npts * 10 + deg)
    for _ in range(5):
        c = rng.normal(size=deg + 1)
        p = lambda t: np.polyval(c[::-1], t)
        hist = np.array([p(-k) for k in range(npts)])  # newest first
        assert abs(w @ hist - p(1.0)) <= 1e-12 * (1 + np.abs(hist).max())


def test_lsq_weights_known_values():
    """Two-point linear and three-point quadratic extrapolation, and the
    six-point quadratic's noise gain (l2 norm of its weights)."""
    assert np.allclose(lsq_weights(2, 1), [2.0, -1.0])
    assert np.allclose(lsq_weights(3, 2), [3.0, -3.0, 1.0])
    w = lsq_weights(6, 2)
    assert np.allclose(w, [1.5, 0.3, -0.4, -0.6, -0.3, 0.5])
    assert np.sqrt((w**2).sum()) < np.sqrt((lsq_weights(3, 2) ** 2).sum()) / 2


def test_lsq_weights_reject_bad_shapes():
    with pytest.raises(ValueError):
        lsq_weights(3, 3)
    with pytest.raises(ValueError):
        lsq_weights(4, 0)


def test_lsq_weights_at_uneven_times():
    """Exact for quadratics at uneven times; equal to the uniform weights at
    uniform times."""
    from fiber3d_amd.learned import lsq_weights_at
    times = np.array([1.0, 0.9, 0.85, 0.7, 0.6, 0.45])   # newest first, uneven steps
    c = np.array([0.3, -1.2, 2.5])
    p = lambda t: c[0] + c[1] * t + c[2] * t * t
    w = lsq_weights_at(times, 1.07, 2)
    assert abs(w @ p(times) - p(1.07)) < 1e-12
    u = lsq_weights_at(0.5 - 0.01 * np.arange(6), 0.51, 2)
    assert np.allclose(u, lsq_weights(6, 2), atol=1e-12)


def test_adaptive_decision_rules():
    """tstep.check_error's accept/reject and step-size rules."""
    from fiber3d_amd.tstep import adaptive_decision
    tol, lo, hi = 1e-2, 1e-5, 1e-2
    assert adaptive_decision(0.5e-2, False, 1e-3, tol, lo, hi) == (True, 1.2e-3)
    assert adaptive_decision(0.95e-2, False, 1e-3, tol, lo, hi) == (True, 1e-3)
    assert adaptive_decision(2e-2, False, 1e-3, tol, lo, hi) == (False, 0.5e-3)
    assert adaptive_decision(0.5e-2, True, 1e-3, tol, lo, hi) == (False, 0.5e-3)
    assert adaptive_decision(0.1e-2, False, 9e-3, tol, lo, hi) == (True, 1e-2)     # capped at dt_max
    assert adaptive_decision(2e-2, False, 1.5e-5, tol, lo, hi) == (True, 1e-5)     # floor: accepted


def test_fibers_reaching_periphery():
    """The reference's periphery-reaching test (tstep_utils.py:343-361) for
    a sphere and an ellipsoid: a fiber whose plus end lies within 5% of the
    periphery's distance, or beyond it, is flagged; one stopping short is
    not."""
    from fiber3d_amd.tstep import fibers_reaching_periphery
    s = np.linspace(0.0, 1.0, 16)[:, None]
    u = np.array([0.6, 0.0, 0.8])
    for a, b, c in [(2.0, 2.0, 2.0), (3.0, 2.5, 2.0)]:
        R = 1.0 / np.sqrt((u[0] / a) ** 2 + (u[1] / b) ** 2 + (u[2] / c) ** 2)  # along u
        fibers = [(R * f - 1.0 + s) * u for f in (0.9, 0.96, 0.999, 1.02)]
        assert fibers_reaching_periphery(fibers, a, b, c) == [1, 2, 3]
        # without the margin only a node on or outside the periphery counts
        assert fibers_reaching_periphery(fibers, a, b, c, margin=0.0) == [3]


class _Sys:
    def __init__(self, dt):
        self.dt = dt


class _Stepper:
    """The parts of TimeStepper the extrapolating predictor reads."""

    def __init__(self, n, fiber_x, dt):
        self.time = 0.0
        self.sys = _Sys(dt)
        self.idx_x = torch.as_tensor(fiber_x, dtype=torch.long).reshape(-1, 3)
        self.n = n


@pytest.mark.parametrize("correct", [False, True])
def test_ring_extrapolation_matches_direct_formula(correct):
    """The predictor's history ring and its one-product extrapolation give,
    on the host, the direct formula: 2 p1 - p2 (at equal steps) for every
    component, and the least-squares quadratic through the last six values
    for the fiber positions; exact for quadratic histories at uneven steps;
    with corrections, the same applied to x + c.  More solutions than the
    ring holds wrap around without changing the result."""
    from fiber3d_amd.learned import LearnedGuess, lsq_weights_at
    n = 40
    fiber_x = np.arange(6, 30)  # 8 nodes x 3 coordinates
    rng = np.random.default_rng(5)
    c0, c1, c2 = rng.normal(size=(3, n))
    cc = rng.normal(size=n) * 1e-3
    x = lambda t: c0 + c1 * t + c2 * t * t
    steps = [0.1, 0.1, 0.07, 0.12, 0.1, 0.09, 0.11, 0.1, 0.05]
    pred = LearnedGuess(train_steps=10**9, use_net=False, correct=correct)
    ts = _Stepper(n, fiber_x, steps[0])
    times = []
    for k, h in enumerate(steps):
        ts.time += h
        times.append(ts.time)
        sol = torch.from_numpy(x(ts.time))
        corr = torch.from_numpy(cc * (k + 1)) if correct else None
        pred.observe(ts, sol, corr)
        ts.sys.dt = steps[(k + 1) % len(steps)]
        if k == 0:
            continue
        got = pred.predict(ts).numpy()
        hist = [x(t) + (cc * (j + 1) if correct else 0.0) for j, t in enumerate(times)]
        t1, t2, tn = times[-1], times[-2], ts.time + ts.sys.dt
        a = (tn - t1) / (t1 - t2)
        want = (1 + a) * hist[-1] - a * hist[-2]
        if len(times) >= 6:
            w = lsq_weights_at(np.array(times[::-1][:6]), tn, 2)
            want[fiber_x] = sum(w[j] * hist[-1 - j][fiber_x] for j in range(6))
            if not correct:
                np.testing.assert_allclose(got[fiber_x], x(tn)[fiber_x], rtol=0, atol=1e-12)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    assert pred.wants_correction == correct
