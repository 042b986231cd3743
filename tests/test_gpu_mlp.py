"""The selector network's HIP kernels against the FP64 oracle
(oracle/mlp_oracle.py): loss and every parameter gradient on random data
(FP32 tolerance), Adam training lowering the loss on a separable problem,
deterministic training, and predictions equal to the oracle's argmax."""
import ctypes as C

import numpy as np
import pytest

from clspmv_amd import selector as S
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d,h,c,n", [(24, 128, 12, 300), (5, 16, 3, 7), (24, 33, 2, 1000)])
def test_loss_and_gradients_match_oracle(ctx, d, h, c, n):
    rng = np.random.default_rng(d * 1000 + n)
    net = S.Mlp(ctx, d, h, c, seed=3)
    p = net.get_params()
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    y = rng.integers(0, c, n).astype(np.int32)
    loss, g = net.loss_grad(x, y)
    loss64, g64 = MO.loss_grad(p, x, y, d, h, c)
    assert abs(loss - loss64) <= 1e-5 * max(1.0, abs(loss64))
    scale = np.max(np.abs(g64))
    assert np.max(np.abs(g - g64)) <= 1e-4 * scale, np.max(np.abs(g - g64)) / scale
    net.free()


def test_training_learns_and_is_deterministic(ctx):
    rng = np.random.default_rng(5)
    d, c, n = 24, 6, 1200
    centers = rng.normal(0, 3, (c, d))
    y = rng.integers(0, c, n).astype(np.int32)
    x = (centers[y] + rng.normal(0, 1, (n, d))).astype(np.float32)
    a = S.Mlp(ctx, d, 64, c, seed=11)
    l0, _ = a.loss_grad(x, y)
    l1 = a.train(x, y, steps=300, lr=3e-3)
    assert l1 < 0.2 * l0, (l0, l1)
    pred = a.predict(x)[1]
    assert np.mean(pred == y) > 0.95
    b = S.Mlp(ctx, d, 64, c, seed=11)
    b.train(x, y, steps=300, lr=3e-3)
    assert np.array_equal(a.get_params().view(np.uint32), b.get_params().view(np.uint32))
    # predictions are the oracle's argmax of the trained parameters
    z64 = MO.forward(a.get_params(), x, d, 64, c)[2]
    assert np.mean(np.argmax(z64, axis=1) == pred) > 0.999
    a.free()
    b.free()


def test_adam_steps_match_oracle_across_calls(ctx):
    """Each GPU Adam step equals the FP64 Adam oracle applied to the GPU's
    own gradient at the GPU's own iterate (within 1e-6 relative), over three
    train() calls of one step each -- the step number continues across
    calls -- and one call of three steps takes bitwise the same steps."""
    rng = np.random.default_rng(21)
    d, h, c, n = 24, 32, 5, 200
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    y = rng.integers(0, c, n).astype(np.int32)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    a = S.Mlp(ctx, d, h, c, seed=7)
    probe = S.Mlp(ctx, d, h, c, seed=7)
    p_prev = a.get_params()
    m = np.zeros(p_prev.size)
    v = np.zeros(p_prev.size)
    for t in (1, 2, 3):
        probe.set_params(p_prev)
        _, g = probe.loss_grad(x, y)
        a.train(x, y, steps=1, lr=lr, beta1=b1, beta2=b2, eps=eps)
        p_gpu = a.get_params()
        p_or, m, v = MO.adam_step(p_prev, g, m, v, t, lr, b1, b2, eps)
        err = np.abs(p_gpu.astype(np.float64) - p_or) / np.maximum(np.abs(p_or), lr)
        assert err.max() <= 1e-6, (t, err.max())
        assert np.abs(p_gpu.astype(np.float64) - p_prev).max() > 0.5 * lr  # the step did move
        p_prev = p_gpu
    b = S.Mlp(ctx, d, h, c, seed=7)
    b.train(x, y, steps=3, lr=lr, beta1=b1, beta2=b2, eps=eps)
    assert np.array_equal(a.get_params().view(np.uint32), b.get_params().view(np.uint32))
    for net in (a, b, probe):
        net.free()
