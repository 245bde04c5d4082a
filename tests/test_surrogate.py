"""Surrogate network (FP32 MLP 5-64-64-2, SiLU) of the hyperparameter search.

CPU: the FP64 oracle's gradient against central finite differences, so the
oracle's backward pass is pinned independently of the kernels.
GPU: the HIP forward, gradient and Adam steps against the oracle:
  forward  normwise 1e-5 relative;
  gradient normwise 1e-4 relative (FP32 sums over 64-wide dot products);
  50 Adam steps: parameters within 1e-3 (normwise) and loss within 1e-3
  relative of the FP64 run from the same start.
"""

import numpy as np
import pytest

from oracle import surrogate as oracle_net
from pypolymlp_amd import capi


def _init(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for shp in oracle_net.SIZES:
        fan_in = shp[1] if len(shp) == 2 else 1
        parts.append(rng.normal(0.0, 1.0 / np.sqrt(fan_in), size=shp).ravel() if len(shp) == 2
                     else 0.1 * rng.normal(size=shp))
    return np.concatenate(parts)


def _data(n=32, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.5, 1.5, size=(n, 5))
    y = np.stack([np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2], np.cos(x[:, 3]) - 0.2 * x[:, 4]], 1)
    return x, y


def test_oracle_gradient_finite_differences():
    p = _init()
    x, y = _data(16)
    _, g = oracle_net.loss_grad(p, x, y)
    rng = np.random.default_rng(2)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (oracle_net.loss_grad(p + e, x, y)[0] - oracle_net.loss_grad(p - e, x, y)[0]) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9)
    assert oracle_net.n_params() == len(p) == 4674


def test_oracle_adam_first_step_closed_form():
    """With zero moments, the bias-corrected first Adam step is
    m_hat = g and v_hat = g^2, so every parameter moves by
    lr * (g / (|g| + eps) + wd * p): about lr * sign(g) wherever |g| >> eps."""
    p = _init(7)
    x, y = _data(16, seed=8)
    _, g = oracle_net.loss_grad(p, x, y)
    lr, wd, eps = 1e-3, 1e-2, 1e-8
    p1, m1, v1, _ = oracle_net.adam(p, np.zeros_like(p), np.zeros_like(p), 0, x, y, 1, lr,
                                    eps=eps, wd=wd)
    np.testing.assert_allclose(p1, p - lr * (g / (np.abs(g) + eps) + wd * p), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v1, 1e-3 * g * g, rtol=1e-12)
    big = np.abs(g) > 1e-4
    np.testing.assert_allclose((p - p1)[big] - lr * wd * p[big], lr * np.sign(g[big]), rtol=1e-4)
    # the step count continues from t0: with t0 = 9 the first update uses the
    # bias corrections 1 - 0.9^10 and 1 - 0.999^10
    p9, _, _, _ = oracle_net.adam(p, np.zeros_like(p), np.zeros_like(p), 9, x, y, 1, lr, eps=eps)
    mh, vh = 0.1 * g / (1 - 0.9 ** 10), 1e-3 * g * g / (1 - 0.999 ** 10)
    np.testing.assert_allclose(p9, p - lr * mh / (np.sqrt(vh) + eps), rtol=0, atol=1e-14)


@pytest.mark.gpu
def test_surrogate_forward_and_gradient_vs_oracle():
    assert capi.surrogate_n_params() == oracle_net.n_params()
    p = _init()
    x, y = _data(32)
    yp = capi.surrogate_predict(p, x)
    yr, _ = oracle_net.forward(p, x)
    assert np.linalg.norm(yp - yr) <= 1e-5 * np.linalg.norm(yr)
    g, loss = capi.surrogate_grad(p, x, y)
    lr, gr = oracle_net.loss_grad(p, x, y)
    assert loss == pytest.approx(lr, rel=1e-5)
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    # predictions in chunks (more than one launch)
    xb = np.random.default_rng(3).uniform(-1, 1, size=(300, 5))
    assert np.linalg.norm(capi.surrogate_predict(p, xb) - oracle_net.forward(p, xb)[0]) <= \
        1e-5 * np.linalg.norm(oracle_net.forward(p, xb)[0])


@pytest.mark.gpu
def test_surrogate_adam_matches_oracle():
    p0 = _init(5)
    x, y = _data(32, seed=6)
    p = p0.astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    loss = capi.surrogate_train(p, m, v, 0, x, y, steps=50, lr=3e-3, weight_decay=1e-2)
    pr, mr, vr, lr_ = oracle_net.adam(p0, np.zeros_like(p0), np.zeros_like(p0), 0, x, y, 50, 3e-3,
                                      wd=1e-2)
    assert np.linalg.norm(p - pr) <= 1e-3 * np.linalg.norm(pr)
    assert loss == pytest.approx(lr_, rel=1e-3)
    # continuing from t0 = 50 reduces the loss further
    loss2 = capi.surrogate_train(p, m, v, 50, x, y, steps=500, lr=3e-3)
    assert loss2 < 0.5 * loss
