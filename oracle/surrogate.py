"""Oracle (test infrastructure only): the surrogate network in FP64 numpy.

Same layout and loss as the HIP kernels (kernels_surrogate.hip): parameter
vector [W1 (64x5) | b1 | W2 (64x64) | b2 | W3 (2x64) | b3], SiLU
activations, loss = mean over samples and outputs of (y - t)^2, Adam with
beta (0.9, 0.999), eps 1e-8, bias correction and decoupled weight decay.
"""

from __future__ import annotations

import numpy as np

NI, NH, NO = 5, 64, 2
SIZES = [(NH, NI), (NH,), (NH, NH), (NH,), (NO, NH), (NO,)]


def unpack(p):
    out, o = [], 0
    for shp in SIZES:
        n = int(np.prod(shp))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(shp))
        o += n
    return out


def n_params():
    return sum(int(np.prod(s)) for s in SIZES)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = x @ W1.T + b1
    h1 = z1 * _sig(z1)
    z2 = h1 @ W2.T + b2
    h2 = z2 * _sig(z2)
    return h2 @ W3.T + b3, (z1, h1, z2, h2)


def loss_grad(p, x, t):
    """(loss, gradient) of mean (y - t)^2."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (z1, h1, z2, h2) = forward(p, x)
    n = len(x)
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / (n * NO)
    gW3, gb3 = dy.T @ h2, dy.sum(0)
    s2 = _sig(z2)
    dz2 = (dy @ W3) * s2 * (1 + z2 * (1 - s2))
    gW2, gb2 = dz2.T @ h1, dz2.sum(0)
    s1 = _sig(z1)
    dz1 = (dz2 @ W2) * s1 * (1 + z1 * (1 - s1))
    gW1, gb1 = dz1.T @ x, dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def adam(p, m, v, t0, x, t, steps, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    p, m, v = (np.array(a, dtype=np.float64) for a in (p, m, v))
    loss = None
    for it in range(steps):
        loss, g = loss_grad(p, x, t)
        k = t0 + it + 1
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * ((m / (1 - b1 ** k)) / (np.sqrt(v / (1 - b2 ** k)) + eps) + wd * p)
    return p, m, v, loss
