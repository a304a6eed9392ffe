"""CPU restatement of the parameter chooser's network (test infrastructure
only): 6 -> 64 -> 64 -> 64 -> 3, ReLU, weighted mean-squared error, Adam, in
float64, with the parameter layout of ducc_amd/csrc/mlp.hip (W1 64x6, b1, W2
64x64, b2, W3 64x64, b3, W4 3x64, b4; rows are output neurons).  Its gradient
is pinned by central differences in the tests."""
import numpy as np

SIZES = ((6, 64), (64, 64), (64, 64), (64, 3))


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for nin, nout in SIZES:
        W = p[o:o + nin * nout].reshape(nout, nin)
        o += nin * nout
        b = p[o:o + nout]
        o += nout
        out.append((W, b))
    assert o == p.size
    return out


def forward(p, X):
    layers = unpack(p)
    a = np.asarray(X, np.float64)
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = a @ W.T + b
        zs.append(z)
        a = np.maximum(z, 0) if li < 3 else z
        acts.append(a)
    return a, zs, acts


def loss_and_grad(p, X, T, w=None):
    w = np.ones(3) if w is None else np.asarray(w, np.float64)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    B = X.shape[0]
    y, zs, acts = forward(p, X)
    r = y - T
    loss = float(np.sum(w * r * r) / (B * 3))
    d = 2 * w * r / (B * 3)
    layers = unpack(p)
    grads = []
    for li in range(3, -1, -1):
        W, b = layers[li]
        gW = d.T @ acts[li]
        gb = d.sum(0)
        grads.append((gW, gb))
        if li > 0:
            d = (d @ W) * (zs[li - 1] > 0)
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)


def adam(p, X, T, steps, lr, wd=0.0, w=None):
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    b1, b2, e = 0.9, 0.999, 1e-8
    for t in range(1, steps + 1):
        _, g = loss_and_grad(p, X, T, w)
        g = g + wd * p
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + e)
    return p
