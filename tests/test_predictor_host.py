"""Host checks of the initial-guess extrapolation weights."""
import numpy as np
import pytest

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
