"""FP64 CPU reference of the learned alias correction's network.

TEST INFRASTRUCTURE ONLY: the checker of the GPU forward and backward
kernels (psh_mlp_forward, psh_mlp_grad) in tests/.  Same layout as the
device parameters: inputs padded to 32, three hidden layers of 64 (ReLU),
one output in a 16-column last layer (powspec_amd.mlp.unpack).  Its own
analytic gradient is pinned against central finite differences
(tests/test_mlp_ref.py)."""
from __future__ import annotations

import numpy as np

from powspec_amd.mlp import OUT, _SHAPES, unpack


def ref_forward(p, X):
    """FP64 forward: (output, cache)."""
    u = {k: v.astype(np.float64) for k, v in unpack(np.asarray(p)).items()}
    x = np.asarray(X, np.float64)
    h1 = np.maximum(x @ u["W1"] + u["b1"], 0)
    h2 = np.maximum(h1 @ u["W2"] + u["b2"], 0)
    h3 = np.maximum(h2 @ u["W3"] + u["b3"], 0)
    y = h3 @ u["W4"] + u["b4"]
    return y[:, 0], (x, h1, h2, h3, u)


def ref_loss(p, X, T):
    y, _ = ref_forward(p, X)
    return float(np.mean((y - T) ** 2))


def ref_grad(p, X, T):
    """FP64 analytic gradient of mean (y - t)^2, in the flat layout."""
    y, (x, h1, h2, h3, u) = ref_forward(p, X)
    B = len(T)
    dy = np.zeros((B, OUT))
    dy[:, 0] = 2 * (y - T) / B
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    d3 = (dy @ u["W4"].T) * (h3 > 0)
    g["W3"], g["b3"] = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ u["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ u["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ d1, d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in _SHAPES])
