"""The LDS x-tiled layout on the CPU (clspmv_xtile_layout_csr: the host
arrays a CSR part uploads for method 4, no device work).

Checked against its own definition, independently of the builder's code:
every nonzero placed exactly once, each row's entries of a column slice
contiguous, tiles of at most 32 chunks covering their slice in order, and
the chunk flags the kernel trusts -- the carry bit exactly where a chunk's
first row continues the previous chunk of its tile, the fast bit only where
no row crosses a lane boundary, the scan plan's steps and row-crossing bits
enough for every run of equal rows.  Then the kernel's chunk reduction is
emulated lane by lane (the fast path's direct adds; the planned segmented
scan with the DPP row shifts, row broadcasts and wave shifts as the
hardware defines them; the carry slots summed per run of equal rows after
the tile) and y must match the FP64 product."""
import ctypes as C

import numpy as np
import pytest

from clspmv_amd import _lib as L
from clspmv_amd import matrix as M
from tests.helpers import rand_coo

CHUNK, XT_COLS, PAD = 128, 8064, 32768
CARRY, FAST, BC31 = 0x8000, 0x4000, 0x2000  # entry 0 of a chunk
BC15 = 0x8000                                # entry 1; bits 13-14: steps - 1
COLMASK = 0x1FFF


def _run_lengths():
    # rows of 1..97 entries in one slice: segments of every length
    rng = np.random.default_rng(16)
    rows, cols = [], []
    for r in range(700):
        n = (7 * r) % 97 + 1
        rows.append(np.full(n, r, np.int32))
        cols.append(np.sort(rng.choice(4000, n, replace=False)).astype(np.int32))
    r, c = np.concatenate(rows), np.concatenate(cols)
    v = rng.standard_normal(r.size).astype(np.float32)
    return M.CooMatrix.from_arrays(r, c, v, 4000, 700)


def _pairs_and_singles():
    # rows of 1-3 entries per slice over a wide x: fast chunks, rows of two
    # entries aligned to a lane, class-1 rows kept off chunk boundaries
    rng = np.random.default_rng(5)
    h, w = 20000, 60000
    n = rng.integers(1, 4, h)
    r = np.repeat(np.arange(h), n).astype(np.int32)
    c = np.concatenate([np.sort(rng.choice(w, k, replace=False)) for k in n]).astype(np.int32)
    v = rng.standard_normal(r.size).astype(np.float32)
    return M.CooMatrix.from_arrays(r, c, v, w, h)


CASES = [
    ("run_lengths", _run_lengths),
    ("zipf", lambda: M.gen_zipf_rows(5000, 1.8, 16.0, 3000, 11)),
    ("uniform_rows", lambda: M.gen_uniform_rows(3000, 17, 3)),
    ("wide_ragged_slices", lambda: rand_coo(5000, 100003, 200000, 4)),
    ("pairs_and_singles", _pairs_and_singles),
    ("empty_rows", lambda: rand_coo(517, 517, 90, 2)),
]


def _layout(coo):
    csr = M.coo2csr(coo)
    lay = L.XtileLayout()
    L.check(L.lib().clspmv_xtile_layout_csr(C.byref(csr.s), C.byref(lay)), "xtile_layout")
    nb, g, nt, nc = lay.nblocks, lay.groups, lay.ntiles, lay.nchunks

    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt) if n else np.zeros(0, dt)

    out = {"nblocks": nb, "groups": g, "blk_row": arr(lay.blk_row, nb + 1, np.int64),
           "blk_tile": arr(lay.blk_tile, nb * g + 1, np.int64), "tile_slice": arr(lay.tile_slice, nt, np.int64),
           "tile_chunk": arr(lay.tile_chunk, nt + 1, np.int64), "idx": arr(lay.idx, nc * CHUNK, np.int64),
           "val": np.ctypeslib.as_array(lay.val, shape=(nc * CHUNK,)).copy() if nc else np.zeros(0, np.float32)}
    L.lib().clspmv_free_xtile_layout(C.byref(lay))
    return out


def _tiles(lay):
    """(workgroup, block, tile, slice, continues, has_carry, chunk0, chunk1)."""
    g = lay["groups"]
    for w in range(len(lay["blk_tile"]) - 1):
        for t in range(lay["blk_tile"][w], lay["blk_tile"][w + 1]):
            ts = lay["tile_slice"][t]
            yield w, w // g, t, ts >> 2, ts & 1, (ts >> 1) & 1, lay["tile_chunk"][t], lay["tile_chunk"][t + 1]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_entries_placed_once_rows_contiguous(name, make):
    coo = make()
    lay = _layout(coo)
    idx, val = lay["idx"], lay["val"]
    rows_out, cols_out, vals_out = [], [], []
    prev_slice = {}
    for w, b, t, sl, cont, _, c0, c1 in _tiles(lay):
        assert 0 < c1 - c0 <= 32
        assert (w % lay["groups"]) == sl // -(-max(1, -(-coo.width // XT_COLS)) // lay["groups"])
        # a continuing tile follows a tile of the same slice in the same workgroup
        assert bool(cont) == (prev_slice.get(w) == sl), (t, cont)
        prev_slice[w] = sl
    for w, b, t, sl, cont, _, c0, c1 in _tiles(lay):
        e = idx[c0 * CHUNK:c1 * CHUNK]
        r = e >> 16
        real = r != PAD
        rows_out.append(lay["blk_row"][b] + r[real])
        cols_out.append(sl * XT_COLS + (e[real] & COLMASK))
        vals_out.append(val[c0 * CHUNK:c1 * CHUNK][real])
        assert np.all(val[c0 * CHUNK:c1 * CHUNK][~real] == 0)
    # each row's entries of a slice are one contiguous run (over the slice's tiles)
    for w in range(len(lay["blk_tile"]) - 1):
        t0, t1 = lay["blk_tile"][w], lay["blk_tile"][w + 1]
        for sl in set(int(s) for s in (lay["tile_slice"][t0:t1] >> 2)):
            ts = [t for t in range(t0, t1) if lay["tile_slice"][t] >> 2 == sl]
            e = np.concatenate([idx[lay["tile_chunk"][t] * CHUNK:lay["tile_chunk"][t + 1] * CHUNK] for t in ts]) >> 16
            e = e[e != PAD]
            starts = np.flatnonzero(np.r_[True, e[1:] != e[:-1]])
            assert len(starts) == len(set(e.tolist())), (name, w, sl)
    got = sorted(zip(np.concatenate(rows_out).tolist(), np.concatenate(cols_out).tolist(),
                     np.concatenate(vals_out).view(np.uint32).tolist())) if rows_out else []
    want = sorted(zip(coo.row_id.tolist(), coo.col_id.tolist(), coo.data.view(np.uint32).tolist()))
    assert got == want


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_chunk_flags_and_scan_plans(name, make):
    lay = _layout(make())
    idx = lay["idx"]
    counts = {"fast": 0, "carry": 0, "scan": 0}
    for w, b, t, sl, cont, tcarry, c0, c1 in _tiles(lay):
        any_carry = False
        for c in range(c0, c1):
            e = idx[c * CHUNK:(c + 1) * CHUNK]
            rows = e >> 16
            head, tail = rows[0::2], rows[1::2]
            carry = c > c0 and rows[0] == (idx[c * CHUNK - 1] >> 16) and rows[0] != PAD
            assert bool(e[0] & CARRY) == carry, (name, c)
            any_carry |= carry
            crossing = (head[1:] != PAD) & (head[1:] == tail[:-1])
            if e[0] & FAST:
                assert not carry and not crossing.any(), (name, c)
                counts["fast"] += 1
                continue
            counts["carry" if carry else "scan"] += 1
            steps = ((e[1] >> 13) & 3) + 1
            same = (tail[1:] != PAD) & (tail[1:] == tail[:-1])  # lane l (>= 1) continues lane l - 1's tail row
            run = longest = 1
            for l in range(1, 64):
                run = run + 1 if (l % 16 and same[l - 1]) else 1
                longest = max(longest, run)
            assert (1 << steps) >= longest, (name, c, steps, longest)
            assert bool(e[1] & BC15) == bool(same[15] or same[47]), (name, c)
            assert bool(e[0] & BC31) == bool(same[31]), (name, c)
        assert bool(tcarry) == any_carry, (name, t)
    assert sum(counts.values()) == len(idx) // CHUNK


# -------------------------------------------------- the kernel's reduction

def _dpp(v, mode, old):
    """v of another lane as GFX9 DPP defines it; lanes with no source (or
    in a row the mode does not write) keep `old`."""
    out = np.array(old, copy=True)
    lane = np.arange(64)
    if mode[0] == "row_shr":
        d = mode[1]
        ok = lane % 16 >= d
        out[ok] = v[lane[ok] - d]
    elif mode[0] == "row_bcast15":  # rows 1 and 3 read lane 15 of the row below
        for row in (1, 3):
            out[row * 16:(row + 1) * 16] = v[row * 16 - 1]
    elif mode[0] == "row_bcast31":  # rows 2 and 3 read lane 31
        out[32:] = v[31]
    elif mode[0] == "wave_shr1":
        out[1:] = v[:-1]
    elif mode[0] == "wave_shl1":
        out[:-1] = v[1:]
    return out


def _chunk(e, v, x_l, y_l):
    """One chunk as a wave reduces it; returns (carry row or -1, carry sum)."""
    f32 = np.float32
    ea, eb = e[0::2], e[1::2]
    ra, rb = ea >> 16, eb >> 16
    pa = (v[0::2] * x_l[ea & COLMASK]).astype(f32)
    pb = (v[1::2] * x_l[eb & COLMASK]).astype(f32)
    brk = ra != rb
    tail_sum = np.where(brk, pb, (pa + pb).astype(f32)).astype(f32)
    if e[0] & FAST:
        for l in range(64):
            if brk[l]:
                y_l[ra[l]] = f32(y_l[ra[l]] + pa[l])
            if rb[l] != PAD:
                y_l[rb[l]] = f32(y_l[rb[l]] + tail_sum[l])
        return -1, f32(0)
    first, carry = ra[0], bool(e[0] & CARRY)
    steps = ((e[1] >> 13) & 3) + 1
    key = rb + 1
    c = tail_sum.copy()

    def step(mode):
        nonlocal c
        k = _dpp(key, mode, np.zeros(64, np.int64))
        u = _dpp(c, mode, np.zeros(64, f32))
        c = np.where(k == key, (c + u).astype(f32), c).astype(f32)

    for s in range(steps):
        step(("row_shr", 1 << s))
    if e[1] & BC15:
        step(("row_bcast15",))
    if e[0] & BC31:
        step(("row_bcast31",))
    prev_tail = _dpp(rb, ("wave_shr1",), rb)
    next_head = _dpp(ra, ("wave_shl1",), np.full(64, -1, np.int64))
    prev_c = _dpp(c, ("wave_shr1",), c)
    cont = (np.arange(64) > 0) & (prev_tail == ra)
    head_sum = (pa + np.where(cont, prev_c, f32(0))).astype(f32)
    crow, cval = -1, f32(0)
    for on, row, s in ((brk, ra, head_sum), (next_head != rb, rb, c)):
        for l in range(64):
            if not on[l] or row[l] == PAD:
                continue
            if carry and row[l] == first:
                crow, cval = first, s[l]
            else:
                y_l[row[l]] = f32(y_l[row[l]] + s[l])
    return crow, cval


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_emulated_chunk_reduction_matches_fp64(name, make):
    coo = make()
    lay = _layout(coo)
    x = M.gen_vector(coo.width, 5)
    g = lay["groups"]
    tiles = list(_tiles(lay))
    y = np.zeros(coo.height, np.float64)
    for w in range(len(lay["blk_tile"]) - 1):
        b = w // g
        r0, r1 = lay["blk_row"][b], lay["blk_row"][b + 1]
        y_l = np.zeros(32768, np.float32)
        for _, _, t, sl, _, tcarry, c0, c1 in tiles[lay["blk_tile"][w]:lay["blk_tile"][w + 1]]:
            x_l = np.zeros(XT_COLS, np.float32)
            part = x[sl * XT_COLS:min(coo.width, (sl + 1) * XT_COLS)]
            x_l[:part.size] = part
            slots = [_chunk(lay["idx"][c * CHUNK:(c + 1) * CHUNK], lay["val"][c * CHUNK:(c + 1) * CHUNK], x_l, y_l)
                     for c in range(c0, c1)]
            if tcarry:  # runs of equal rows over the slots, in chunk order, added once
                i = 0
                while i < len(slots):
                    r, s = slots[i]
                    j = i + 1
                    while r >= 0 and j < len(slots) and slots[j][0] == r:
                        s = np.float32(s + slots[j][1])
                        j += 1
                    if r >= 0:
                        y_l[r] = np.float32(y_l[r] + s)
                    i = j
        y[r0:r1] += y_l[:r1 - r0]
    ref = np.zeros(coo.height, np.float64)
    np.add.at(ref, coo.row_id, coo.data.astype(np.float64) * x[coo.col_id].astype(np.float64))
    absref = np.zeros(coo.height, np.float64)
    np.add.at(absref, coo.row_id, np.abs(coo.data.astype(np.float64) * x[coo.col_id].astype(np.float64)))
    # an FP32 sum of n rounded products: |err| <= gamma_(n+1) * sum |a x|,
    # gamma_k = k u / (1 - k u), u = 2^-24 (y starts at 0 and takes the
    # block's sum once)
    n = np.bincount(coo.row_id, minlength=coo.height).astype(np.float64) + 2
    u = 2.0 ** -24
    tol = n * u / (1 - n * u) * absref
    assert np.all(np.abs(y - ref) <= tol), (name, float(np.max(np.abs(y - ref) - tol)))
