"""Pins of the FP64 network oracle (CPU only): its gradient equals central
finite differences of its loss, and at zero weights the loss is log(C) with
the closed-form output-bias gradient mean(p - onehot) = 1/C - freq(class)."""
import numpy as np

from oracle import mlp_oracle as MO


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, h, c, n = 5, 7, 4, 9
    np_ = d * h + h + h * h + h + h * c + c
    p = rng.normal(0, 0.5, np_)
    x = rng.normal(0, 1, (n, d))
    y = rng.integers(0, c, n)
    _, g = MO.loss_grad(p, x, y, d, h, c)
    eps = 1e-6
    for i in rng.choice(np_, 40, replace=False):
        e = np.zeros(np_)
        e[i] = eps
        fd = (MO.loss_only(p + e, x, y, d, h, c) - MO.loss_only(p - e, x, y, d, h, c)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_zero_weights_closed_form():
    d, h, c, n = 3, 4, 5, 10
    np_ = d * h + h + h * h + h + h * c + c
    x = np.ones((n, d))
    y = np.array([0, 1, 1, 2, 2, 2, 3, 3, 3, 3])
    loss, g = MO.loss_grad(np.zeros(np_), x, y, d, h, c)
    assert abs(loss - np.log(c)) < 1e-12
    gb3 = g[-c:]
    freq = np.bincount(y, minlength=c) / n
    assert np.allclose(gb3, 1.0 / c - freq)


def test_adam_constant_gradient_closed_form():
    # with a constant gradient g the bias-corrected moments are exactly g and
    # g^2, so every step moves each parameter by lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 1.0, 1.0, 1.0])
    m = np.zeros(4)
    v = np.zeros(4)
    lr, eps = 0.01, 1e-8
    for t in range(1, 6):
        p_new, m, v = MO.adam_step(p, g, m, v, t, lr, 0.9, 0.999, eps)
        assert np.allclose(p_new - p, -lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15), t
        p = p_new
    assert p[3] == 1.0  # zero gradient from zero moments: unchanged


def test_adam_worked_second_step():
    # g1 = 1, g2 = -1, b1 = 0.9, b2 = 0.999 (by hand): m2 = 0.09 - 0.1 = -0.01,
    # m2 / (1 - 0.81) = -0.0526315789..., v2 = 0.000999 + 0.001 = 0.001999 and
    # v2 / (1 - 0.998001) = 1, so step 2 moves p by +0.0526315789 * lr = 0.01 / 1.9
    p, m, v = MO.adam_step(np.zeros(1), np.ones(1), np.zeros(1), np.zeros(1), 1, 0.1, 0.9, 0.999, 0.0)
    assert np.allclose(p, [-0.1], rtol=1e-12)
    p2, m, v = MO.adam_step(p, -np.ones(1), m, v, 2, 0.1, 0.9, 0.999, 0.0)
    assert np.allclose(m, [-0.01], rtol=1e-12) and np.allclose(v, [0.001999], rtol=1e-12)
    assert np.allclose(p2 - p, [0.01 / 1.9], rtol=1e-10)
