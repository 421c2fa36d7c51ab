"""TEST INFRASTRUCTURE ONLY: a float64 numpy statement of the learned-guess
network (three SiLU hidden layers, linear output, masked mean-squared
error), used to check the gradients of the HIP forward/backward kernels.

The reference has no learned guess, so nothing here restates reference code;
the statement itself is pinned by central finite differences
(tests/test_oracle.py).  Parameter order is the library's flat vector:
W1 b1 W2 b2 W3 b3 W4 b4, each W row-major (out x in).
"""
from __future__ import annotations

import numpy as np


def unpack(p, n_in, width, n_out):
    p = np.asarray(p, dtype=np.float64)
    dims = [(n_in, width), (width, width), (width, width), (width, n_out)]
    out, o = [], 0
    for fin, fout in dims:
        W = p[o:o + fin * fout].reshape(fout, fin)
        o += fin * fout
        b = p[o:o + fout]
        o += fout
        out.append((W, b))
    assert o == p.size
    return out


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, X, n_in, width, n_out):
    layers = unpack(p, n_in, width, n_out)
    h = np.asarray(X, dtype=np.float64)
    for W, b in layers[:3]:
        z = h @ W.T + b
        h = z * _sig(z)
    W, b = layers[3]
    return h @ W.T + b


def loss_grad(p, X, T, M, n_in, width, n_out):
    """sum(M (Y - T)^2) / sum(M) and its gradient (flat, library order)."""
    layers = unpack(p, n_in, width, n_out)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    w = max(M.sum(), 1.0)
    hs, zs = [X], []
    h = X
    for W, b in layers[:3]:
        z = h @ W.T + b
        zs.append(z)
        h = z * _sig(z)
        hs.append(h)
    W4, b4 = layers[3]
    Y = h @ W4.T + b4
    d = (Y - T) * M
    loss = float(np.sum(d * d) / w)
    dY = 2.0 * d / w
    grads = [None] * 4
    g = dY
    for l in (3, 2, 1, 0):
        W, _ = layers[l]
        grads[l] = (g.T @ hs[l], g.sum(axis=0))
        if l > 0:
            gh = g @ W
            z = zs[l - 1]
            s = _sig(z)
            g = gh * s * (1.0 + z * (1.0 - s))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat
