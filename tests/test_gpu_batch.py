"""GPU parity of the batched multi-right-hand-side SpMV: every right-hand
side of every matrix of a mixed batch (k in {1, 4, 16}; stencil, uniform,
power-law and block rows; empty rows and an empty matrix; long rows cut into
chunks) against the CPU oracle with the FP32 parity bound; tile shapes at
their limits; and graph replays, each bitwise equal to the eager run."""
import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import batch as B
from clspmv_amd import matrix as M
from oracle import oracle as O
from tests.helpers import rand_coo

pytestmark = pytest.mark.gpu


def _mats():
    return [M.gen_stencil27(12, 1), M.gen_uniform_rows(3000, 17, 2), M.gen_zipf_rows(8000, 1.8, 16.0, 6000, 3),
            M.gen_block_rows(700, 3, 5, 40, 4), rand_coo(517, 601, 90, 5), rand_coo(40, 33, 0, 6),
            rand_coo(2, 70000, 60000, 7)]


def _check_all(b, coos, xs, y0s, what):
    for i, coo in enumerate(coos):
        y = b.get_y(i)
        for j in range(b.ks[i]):
            rep = O.parity_report(y[:, j], coo, xs[i][:, j], y0s[i][:, j])
            assert rep["ok"], (what, i, j, rep)


def _setup(ctx, ks, tiles=None, seed=0):
    coos = _mats()
    csrs = [M.coo2csr(c) for c in coos]
    ks = [ks[i % len(ks)] for i in range(len(coos))]
    b = B.Batch(ctx, csrs, ks, tiles)
    rng = np.random.default_rng(seed)
    xs, y0s = [], []
    for i, c in enumerate(coos):
        xs.append(rng.uniform(-1, 1, (c.width, ks[i])).astype(np.float32))
        y0s.append(rng.uniform(-1, 1, (c.height, ks[i])).astype(np.float32))
        b.set_x(i, xs[-1])
        b.set_y(i, y0s[-1])
    return b, coos, xs, y0s


@pytest.mark.parametrize("ks", [[1], [4], [16], [1, 4, 16]], ids=["k1", "k4", "k16", "mixed"])
def test_batch_matches_oracle(ctx, ks):
    b, coos, xs, y0s = _setup(ctx, ks)
    b.run(1)
    _check_all(b, coos, xs, y0s, f"ks={ks}")
    b.free()


@pytest.mark.parametrize("tile", [(1, 1, 1), (4096, 512, 512), (7, 64, 3), (64, 512, 256)])
def test_batch_tiles(ctx, tile):
    # (1, 1, 1): every row its own item and every nonempty row long
    n = len(_mats())
    b, coos, xs, y0s = _setup(ctx, [16, 1, 4], [tile] * n, seed=1)
    b.run(1)
    _check_all(b, coos, xs, y0s, f"tile={tile}")
    b.free()


def test_batch_rejects_bad_arguments(ctx):
    csr = M.coo2csr(rand_coo(10, 10, 20, 1))
    with pytest.raises(L.ClspmvError):
        B.Batch(ctx, [csr], [3])
    with pytest.raises(L.ClspmvError):
        B.Batch(ctx, [csr], [4], [(4, 1024, 4)])


def test_batch_graph_replays_match_eager(ctx):
    b, coos, xs, y0s = _setup(ctx, [1, 4, 16], seed=2)
    b.run(1)
    eager = [b.get_y(i).copy() for i in range(len(coos))]
    _check_all(b, coos, xs, y0s, "eager")
    b.capture()
    for rep in range(4):
        for i in range(len(coos)):
            b.set_y(i, y0s[i])  # eager work and a sync between replays
        ctx.sync()
        b.run(1)
        for i in range(len(coos)):
            assert np.array_equal(b.get_y(i).view(np.uint32), eager[i].view(np.uint32)), (rep, i)
    b.free()
