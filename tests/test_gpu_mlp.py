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
