"""Learned parameter chooser (host side).

The network (6 -> 64 -> 64 -> 64 -> 3, ReLU) maps a request's features
(ndim, mean log2 grid size, log10 M, -log10 eps, fp64 flag, type-1 flag) to
offsets of (oversampling factor, support W, log2 tile) from ducc0's
analytic choice and from a fixed tile anchor (the tiles of the training set:
2^9 1-D, 2^5 / 2^4 for 2-D FP32 / FP64, 2^3 3-D); the library adds them to
the analytic setting and maps the sum onto the kernel table, so an output
near zero reproduces the analytic kernel and oversampling (and the anchor
tile, which differs from the library's 2-D FP64 default of 2^5).  Its forward pass, gradient and
Adam training run as HIP kernels in libducc_amd.so (mlp.hip).  Trained
weights and normalisation live in ducc_amd/data/chooser_weights.json; without
that file `chooser="learned"` falls back to ducc0's analytic choice.
"""
import ctypes as ct
import json
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "chooser_weights.json")
NIN, NHID, NOUT = 6, 64, 3
_loaded = False


def available():
    return os.path.exists(WEIGHTS)


def nparam():
    return int(_lib.lib().dmx_mlp_nparam())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def init_params(seed=0):
    """He-initialised weights in the library's flat layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in ((NIN, NHID), (NHID, NHID), (NHID, NHID), (NHID, NOUT)):
        parts.append(rng.normal(0, np.sqrt(2.0 / nin), (nout, nin)).ravel())
        parts.append(np.zeros(nout))
    p = _f32(np.concatenate(parts))
    assert p.size == nparam()
    return p


def forward(params, X):
    X = _f32(X)
    B = X.shape[0]
    Y = np.zeros((B, NOUT), np.float32)
    _lib.check(_lib.lib().dmx_mlp_forward(_f32(params).ctypes.data,
                                          X.ctypes.data, B, Y.ctypes.data))
    return Y


def grad(params, X, T, w=None):
    X, T = _f32(X), _f32(T)
    B = X.shape[0]
    w = _f32(np.ones(NOUT) if w is None else w)
    G = np.zeros(nparam(), np.float32)
    loss = np.zeros(1, np.float32)
    _lib.check(_lib.lib().dmx_mlp_grad(_f32(params).ctypes.data,
                                       X.ctypes.data, T.ctypes.data,
                                       w.ctypes.data, B, G.ctypes.data,
                                       loss.ctypes.data))
    return G, float(loss[0])


def train(params, X, T, w=None, steps=2000, lr=1e-3, weight_decay=0.0):
    """Full-batch Adam on the GPU; returns (params, losses)."""
    P = _f32(params).copy()
    X, T = _f32(X), _f32(T)
    w = _f32(np.ones(NOUT) if w is None else w)
    losses = np.zeros(max(steps, 1), np.float32)
    _lib.check(_lib.lib().dmx_mlp_train(P.ctypes.data, X.ctypes.data,
                                        T.ctypes.data, w.ctypes.data,
                                        X.shape[0], int(steps),
                                        ct.c_float(lr),
                                        ct.c_float(weight_decay),
                                        losses.ctypes.data))
    return P, losses[:steps]


def features(shape, npoints, eps, fp64, nu2u):
    f = np.zeros(NIN)
    shp = np.asarray(shape, np.int64)
    _lib.check(_lib.lib().dmx_chooser_features(len(shape), shp.ctypes.data,
                                               int(npoints), float(eps),
                                               int(bool(fp64)),
                                               int(bool(nu2u)), f.ctypes.data))
    return f


def install(params, norm):
    global _loaded
    norm = np.ascontiguousarray(norm, np.float64)
    assert norm.size == 18
    _lib.check(_lib.lib().dmx_chooser_set(_f32(params).ctypes.data,
                                          norm.ctypes.data))
    _loaded = True


def load(path=WEIGHTS):
    d = json.load(open(path))
    install(np.asarray(d["params"], np.float32),
            np.asarray(d["norm"], np.float64))
    return d


def ensure_loaded():
    if not _loaded and available():
        load()


def choose(shape, npoints, eps, values_f64, nu2u, chooser="analytic",
           sigma_min=1.2, sigma_max=2.51, precision="auto"):
    """Parameters a call would use (runs the network on the GPU for
    chooser="learned"), without running the transform."""
    opt = _lib.options(precision=precision, chooser=chooser)
    info = _lib.DmxPlanInfo()
    shp = np.asarray(shape, np.int64)
    _lib.check(_lib.lib().dmx_choose(len(shape), shp.ctypes.data,
                                     int(npoints), float(eps),
                                     int(bool(values_f64)), int(bool(nu2u)),
                                     float(sigma_min), float(sigma_max),
                                     ct.byref(opt), ct.byref(info)))
    return info.as_dict()
