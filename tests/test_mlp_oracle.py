"""Pin the network oracle (oracle/mlp.py) before it checks the HIP
kernels: analytic gradient vs central differences in float64."""
import numpy as np

from oracle import mlp as OM


def test_oracle_gradient_central_differences():
    rng = np.random.default_rng(0)
    n = sum(a * b + b for a, b in OM.SIZES)
    p = rng.normal(0, 0.3, n)
    X = rng.normal(0, 1, (17, 6))
    T = rng.normal(0, 1, (17, 3))
    w = np.array([1.0, 0.5, 2.0])
    _, g = OM.loss_and_grad(p, X, T, w)
    idx = rng.choice(n, 60, replace=False)
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pm = p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (OM.loss_and_grad(pp, X, T, w)[0] -
              OM.loss_and_grad(pm, X, T, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_oracle_adam_decreases_loss():
    rng = np.random.default_rng(1)
    n = sum(a * b + b for a, b in OM.SIZES)
    p = rng.normal(0, 0.2, n)
    X = rng.normal(0, 1, (64, 6))
    T = np.stack([X[:, 0] + X[:, 1], X[:, 2] * 0.5, -X[:, 3]], 1)
    l0 = OM.loss_and_grad(p, X, T)[0]
    p2 = OM.adam(p, X, T, 200, 1e-2)
    assert OM.loss_and_grad(p2, X, T)[0] < 0.2 * l0
