"""The MLP oracle's gradients against central finite differences of the loss."""
import numpy as np

from oracle import mlp as OM


def test_mlp_gradients_finite_difference():
    rng = np.random.default_rng(0)
    F, W, O, M = 5, 16, 2, 40
    params = [rng.normal(size=(F, W)) / 2, rng.normal(size=(W, W)) / 4, rng.normal(size=(W, O)) / 4,
              rng.normal(size=W) * 0.1, rng.normal(size=W) * 0.1, rng.normal(size=O) * 0.1]
    X = rng.normal(size=(M, F))
    T = rng.normal(size=(M, O))
    g = OM.grads(params, X, T)
    h = 1e-6
    for pi, p in enumerate(params):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = OM.loss(params, X, T)
            flat[idx] = old - h
            lm = OM.loss(params, X, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[pi].reshape(-1)[idx]) <= 1e-6 * max(1.0, abs(fd)), (pi, idx)
