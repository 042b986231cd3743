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
