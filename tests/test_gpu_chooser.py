"""The chooser network's HIP kernels against the CPU oracle
(oracle/mlp.py), and the learned chooser's behaviour: valid, accurate
choices and the analytic fallback."""
import json
import os

import numpy as np
import pytest

from oracle import core as O
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_vs_oracle(gpu):
    from ducc_amd import chooser
    rng = np.random.default_rng(0)
    p = chooser.init_params(3)
    X = rng.normal(0, 1, (50, 6)).astype(np.float32)
    T = rng.normal(0, 1, (50, 3)).astype(np.float32)
    w = np.array([1.0, 0.5, 2.0], np.float32)
    y = chooser.forward(p, X)
    yref = OM.forward(p, X)[0]
    assert np.allclose(y, yref, rtol=1e-4, atol=1e-5)
    g, loss = chooser.grad(p, X, T, w)
    lref, gref = OM.loss_and_grad(p, X, T, w)
    assert abs(loss - lref) <= 1e-5 * lref
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)
    # deterministic: same inputs, same bits
    g2, _ = chooser.grad(p, X, T, w)
    assert np.array_equal(g, g2)


def test_training_matches_oracle_adam(gpu):
    from ducc_amd import chooser
    rng = np.random.default_rng(1)
    p = chooser.init_params(4)
    X = rng.normal(0, 1, (64, 6)).astype(np.float32)
    T = np.stack([X[:, 0] + X[:, 1], X[:, 2] * 0.5, -X[:, 3]], 1)
    T = T.astype(np.float32)
    P, losses = chooser.train(p, X, T, steps=50, lr=1e-2)
    Pref = OM.adam(p, X, T, 50, 1e-2)
    assert np.linalg.norm(P - Pref) <= 1e-3 * np.linalg.norm(Pref)
    assert losses[-1] < 0.5 * losses[0]


def test_learned_choice_is_valid_and_accurate(gpu):
    """With trained weights installed, learned choices are tabulated kernels
    reaching eps and the results meet the reference tolerance."""
    from ducc_amd import chooser, nufft
    if not chooser.available():
        pytest.skip("no trained weights committed yet")
    chooser.load()
    rng = np.random.default_rng(7)
    for shape, M, eps in (((300, 200), 50000, 1e-5), ((4000,), 20000, 1e-8),
                          ((40, 30, 20), 30000, 1e-4)):
        nd = len(shape)
        c = rng.uniform(-np.pi, np.pi, (M, nd))
        f64 = eps < 1e-6
        dt = np.complex128 if f64 else np.complex64
        p = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(dt)
        info = {}
        out = np.empty(shape, dt)
        nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=out,
                   chooser="learned", info=info)
        assert info["eps_tab"] <= eps
        sel = rng.choice(int(np.prod(shape)), 64, replace=False)
        ref = O.nudft_type1(c, p, shape, True, modes=sel)
        mul = {1: 4, 2: 2, 3: 1}[nd]
        assert O.l2error(out.ravel()[sel], ref) <= mul * eps


def test_network_choice_and_fallback(gpu):
    """A network whose output is fixed (zero weights: outputs = the target
    means) asks for the analytic setting with one more support point and a
    tile of half the default side: the request uses the tabulated kernel of
    that support nearest the analytic oversampling factor.  When the FP64
    spot-check is forced to fail from the host, the call takes ducc0's
    analytic choice instead, and the result still meets eps."""
    from ducc_amd import chooser, nufft
    rng = np.random.default_rng(8)
    shape, M, eps = (64, 64), 5000, 1e-3
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    ref = O.nudft_type1(c, p, shape, True)
    an = {}
    out = np.empty(shape, np.complex128)
    nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=out,
               chooser="analytic", info=an)
    params = np.zeros(chooser.nparam(), np.float32)
    norm = np.concatenate([np.zeros(6), np.ones(6), [0.0, 1.0, -1.0],
                           np.ones(3)])
    chooser.install(params, norm)
    try:
        info = {}
        nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=out,
                   chooser="learned", info=info)
        assert info["chooser_used"] and info["W"] == an["W"] + 1
        assert info["log2tile"] == 4 - 1   # 2-D FP64 tile anchor: 2^4
        assert info["eps_tab"] <= eps
        assert O.l2error(out, ref) <= 2 * eps
        info2 = {}
        shape2 = (64, 62)   # a new request signature: spot-checked again
        out2 = np.empty(shape2, np.complex128)
        nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=out2,
                   chooser="learned", info=info2, spot_check="fail")
        assert not info2["chooser_used"]
        k, nov = O.find_params(eps, shape2, M, True, False, vlen=8,
                               nthreads=8)
        assert info2["kernel_idx"] == k and tuple(info2["nover"]) == nov
        ref2 = O.nudft_type1(c, p, shape2, True)
        assert O.l2error(out2, ref2) <= 2 * eps
    finally:
        if chooser.available():
            chooser.load()
