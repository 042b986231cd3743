"""FP64 numpy restatement of the selector network (TEST INFRASTRUCTURE ONLY:
used by tests/ as the checker of the HIP training kernels, never by the
library).  Same parameter layout as include/clspmv_mlp.h: W1 (in x hidden),
b1, W2 (hidden x hidden), b2, W3 (hidden x classes), b3; ReLU hidden layers,
softmax cross-entropy averaged over the samples, Adam.

Pinned by: central finite differences of the loss (gradients), the
closed form of the softmax cross-entropy gradient at zero weights
(p = 1/C for every class), and, for the Adam step, closed forms of
Kingma & Ba's Algorithm 1 (arXiv:1412.6980): with a constant gradient g
the bias-corrected moments are exactly g and g^2 at every step, so each
step moves a parameter by lr * g / (|g| + eps) (about lr * sign(g)); a
step with g = 0 from zero moments leaves it unchanged."""
import numpy as np


def split(p, d, h, c):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((d, h), (h,), (h, h), (h,), (h, c), (c,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, x, d, h, c):
    W1, b1, W2, b2, W3, b3 = split(p, d, h, c)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h1, h2, h2 @ W3 + b3


def loss_grad(p, x, y, d, h, c):
    W1, b1, W2, b2, W3, b3 = split(p, d, h, c)
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    h1, h2, z = forward(p, x, d, h, c)
    zm = z - z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(zm).sum(axis=1))
    loss = float(np.mean(lse - zm[np.arange(n), y]))
    pr = np.exp(zm - lse[:, None])
    dz = pr.copy()
    dz[np.arange(n), y] -= 1.0
    dz /= n
    gW3 = h2.T @ dz
    gb3 = dz.sum(axis=0)
    dh2 = (dz @ W3.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def loss_only(p, x, y, d, h, c):
    n = x.shape[0]
    z = forward(p, x, d, h, c)[2]
    zm = z - z.max(axis=1, keepdims=True)
    return float(np.mean(np.log(np.exp(zm).sum(axis=1)) - zm[np.arange(n), y]))


def adam_step(p, g, m, v, t, lr, b1, b2, eps):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in FP64: returns the
    new (p, m, v) for step number t (1-based)."""
    p = np.asarray(p, np.float64)
    g = np.asarray(g, np.float64)
    m = b1 * np.asarray(m, np.float64) + (1.0 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
