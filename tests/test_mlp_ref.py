"""FP64 reference of the alias-correction network: analytic gradients
against central finite differences, and the padding stays inert."""
import numpy as np

from oracle import mlp_ref as ref
from powspec_amd import mlp


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp.init_params(3).astype(np.float64)
    X = np.zeros((40, mlp.IN))
    X[:, : mlp.NFEAT] = rng.normal(0, 1, (40, mlp.NFEAT))
    T = rng.normal(0, 0.1, 40)
    g = ref.ref_grad(p, X, T)
    live = np.nonzero(mlp.init_params(3))[0]
    pick = rng.choice(live, 60, replace=False)
    h = 1e-6
    for i in pick:
        e = np.zeros_like(p)
        e[i] = h
        fd = (ref.ref_loss(p + e, X, T) - ref.ref_loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (abs(g[i]) + 1e-3), (i, fd, g[i])


def test_padding_gets_no_gradient():
    rng = np.random.default_rng(2)
    p = mlp.init_params(4).astype(np.float64)
    X = np.zeros((30, mlp.IN))
    X[:, : mlp.NFEAT] = rng.normal(0, 1, (30, mlp.NFEAT))
    g = mlp.unpack(ref.ref_grad(p, X, rng.normal(0, 1, 30)))
    assert np.all(g["W1"][mlp.NFEAT:] == 0)
    assert np.all(g["W4"][:, 1:] == 0) and np.all(g["b4"][1:] == 0)
    assert 9900 <= sum(np.count_nonzero(v) for v in mlp.unpack(mlp.init_params(4)).values()) + 64 * 3 + 1 <= 10100
