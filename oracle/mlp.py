"""Oracle: FP64 forward/backward of the per-node MLP (3 layers, ReLU).

No reference counterpart (the learned initial guess is new).  Pinned by
tests/test_oracle_mlp.py against central finite differences of the loss.
"""
import numpy as np


def forward(params, X):
    W1, W2, W3, b1, b2, b3 = params
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0.0)
    Y = H2 @ W3 + b3
    return Y, (X, Z1, H1, Z2, H2)


def loss(params, X, T):
    Y, _ = forward(params, X)
    return np.mean((Y - T) ** 2)


def grads(params, X, T):
    """Gradients of mean((Y - T)^2) wrt (W1, W2, W3, b1, b2, b3)."""
    W1, W2, W3, b1, b2, b3 = params
    Y, (X, Z1, H1, Z2, H2) = forward(params, X)
    dY = 2.0 * (Y - T) / Y.size
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * (Z2 > 0)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    return [gW1, gW2, gW3, gb1, gb2, gb3]
