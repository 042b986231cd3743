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


def _block_cases():
    # (coo, block dim): dense blocks, ragged edges (rows and columns not a
    # multiple of the block), uneven blocks per block row and empty block
    # rows, and a matrix whose entries do not fill their blocks
    blocks3 = M.gen_block_rows(700, 3, 9, 40, 11)
    blocks4 = M.gen_block_rows(300, 4, 6, 30, 12)
    blocks2 = M.gen_block_rows(900, 2, 7, 60, 13)
    keep = (blocks3.row_id < 600) | (blocks3.row_id >= 900)
    holes = M.CooMatrix.from_arrays(blocks3.row_id[keep].copy(), blocks3.col_id[keep].copy(),
                                    blocks3.data[keep].copy(), blocks3.width, blocks3.height)
    return [(blocks3, 3), (blocks4, 4), (blocks2, 2), (holes, 3), (rand_coo(301, 299, 3000, 14), 3),
            (rand_coo(37, 41, 300, 15), 4)]


@pytest.mark.parametrize("k", [4, 16])
def test_batch_bsr_blocks_on_matrix_cores(ctx, k):
    # BSR entries of a batch (block items on v_mfma_f32_4x4x1f32) beside
    # CSR entries, every right-hand side against the oracle; graph replays
    # bitwise equal to the eager run
    cases = _block_cases()
    coos = [c for c, _ in cases] + [M.gen_zipf_rows(5000, 1.8, 16.0, 3000, 16)]
    mats = [M.coo2bsr(c, b, b) for c, b in cases] + [M.coo2csr(coos[-1])]
    ks = [k] * len(mats)
    b = B.Batch(ctx, mats, ks)
    rng = np.random.default_rng(k)
    xs, y0s = [], []
    for i, c in enumerate(coos):
        xs.append(rng.uniform(-1, 1, (c.width, k)).astype(np.float32))
        y0s.append(rng.uniform(-1, 1, (c.height, k)).astype(np.float32))
        b.set_x(i, xs[-1])
        b.set_y(i, y0s[-1])
    b.run(1)
    _check_all(b, coos, xs, y0s, f"bsr k={k}")
    eager = [b.get_y(i) for i in range(len(coos))]
    st = b.stats()
    assert st["nnz"] == sum(c.nnz for c in coos)
    b.capture()
    for rep in range(3):
        for i in range(len(coos)):
            b.set_y(i, y0s[i])
        b.run(1)
        for i in range(len(coos)):
            assert np.array_equal(b.get_y(i).view(np.uint32), eager[i].view(np.uint32)), (rep, i)
    b.free()


def test_batch_bsr_refuses_k1_and_bad_blocks(ctx):
    coo = M.gen_block_rows(50, 3, 4, 10, 17)
    for mat, k in ((M.coo2bsr(coo, 3, 3), 1), (M.coo2bsr(coo, 3, 2), 4), (M.coo2bsr(coo, 5, 5), 4)):
        with pytest.raises(L.ClspmvError):
            B.Batch(ctx, [mat], [k])
    assert b"batch" in L.lib().clspmv_last_error()
