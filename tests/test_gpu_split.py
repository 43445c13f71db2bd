"""GPU tests of the split operator (mgv_split) against the oracle's split
(oracle/mgv_oracle.c orc_split; src/operators/split.cu:49-118): the rows of
each parent must be the oracle's multiset (the order inside a parent is
unspecified), for one to four columns, with and without a filter, one, two
and more parents, ragged and empty vectors, several calls appending to the
same cursors, a small explicit grid, outputs that overflow, bad arguments,
and a 2^25-row vector checked through per-parent checksums."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

VP = C.c_void_p


def canon(cols):
    """The rows of a column list in a canonical (sorted) order."""
    m = np.stack([np.asarray(c, np.int64) for c in cols]) if len(cols) else np.zeros((0, 0), np.int64)
    if m.shape[1] == 0:
        return m
    return m[:, np.lexsort(m[::-1])]


class Split:
    """Device outputs and cursors of one split with `nparents` parents."""

    def __init__(self, nparents, ncols, cap):
        self.nparents, self.ncols, self.cap = nparents, ncols, cap
        # one guard row past `cap` in every output, to see that nothing is written there
        self.out = [mgv.device_array(np.full(cap + 1, -7, np.int32)) for _ in range(nparents * ncols)]
        self.counts = mgv.device_array(np.zeros(nparents, np.uint64))

    def run(self, cols, pred=None, bound=0, grid=0):
        n = len(cols[0])
        dcols = [mgv.device_array(np.ascontiguousarray(c, np.int32)) for c in cols]
        dpred = mgv.device_array(np.ascontiguousarray(pred, np.int32)) if pred is not None else None
        cp = (VP * len(dcols))(*[d.ptr for d in dcols])
        op = (VP * len(self.out))(*[o.ptr for o in self.out])
        mgv.call("mgv_split", dpred.vp if dpred else None, bound, cp, len(dcols), n, self.nparents, op, self.cap,
                 self.counts.vp, mgv.launch(grid, 1024), None)
        mgv.sync()

    def cursors(self):
        return self.counts.download(np.uint64, self.nparents).astype(np.int64)

    def parent(self, p):
        k = min(int(self.cursors()[p]), self.cap)
        return [self.out[p * self.ncols + c].download(np.int32, k) for c in range(self.ncols)]

    def guard(self, p, c):
        return int(self.out[p * self.ncols + c].download(np.int32, self.cap + 1)[self.cap])


def seeded_cols(orc, n, ncols, cfg):
    cols = [orc.gen_bits(n, orc.seed(cfg, c), 9, 32) for c in range(ncols)]
    return cols


@pytest.mark.parametrize("n", [0, 1, 63, 64, 8191, 8192, 8193, 100000, (1 << 20) + 37])
@pytest.mark.parametrize("ncols,nparents,filtered", [(1, 2, True), (2, 2, False), (4, 3, True), (3, 1, False)])
def test_split_matches_oracle(gpu, orc, n, ncols, nparents, filtered):
    cols = seeded_cols(orc, n, ncols, 20 + ncols)
    pred = orc.gen_range(n, orc.seed(20, 9), 9, 1000) if filtered else None
    bound = 499
    want = orc.split(cols, nparents, pred=pred, bound=bound)
    s = Split(nparents, ncols, max(n, 1))
    s.run(cols, pred, bound)
    got_counts = s.cursors()
    for p in range(nparents):
        assert got_counts[p] == len(want[p][0])
        assert np.array_equal(canon(s.parent(p)), canon(want[p]))


def test_split_appends_to_the_cursors(gpu, orc):
    """Three vectors through the same cursors: each parent's output holds the
    rows of all three, as the reference's staging carries over between
    consume_warp calls."""
    n = [5000, 70000, 1]
    vecs = [seeded_cols(orc, k, 2, 30 + i) for i, k in enumerate(n)]
    s = Split(2, 2, sum(n))
    for v in vecs:
        s.run(v)
    allc = [np.concatenate([v[c] for v in vecs]) for c in range(2)]
    want = orc.split(allc, 2)
    assert s.cursors().tolist() == [len(want[0][0]), len(want[1][0])]
    for p in range(2):
        assert np.array_equal(canon(s.parent(p)), canon(want[p]))


@pytest.mark.parametrize("n", [16383, 16384, 16385, 16384 * 5 + 1])
@pytest.mark.parametrize("filtered", [False, True])
def test_split_one_column_tile_edges(gpu, orc, n, filtered):
    """One column is split in tiles of 16384 rows (512 threads x 32 rows):
    inputs just below, at and above a tile, and one row past five tiles."""
    cols = seeded_cols(orc, n, 1, 45)
    pred = orc.gen_range(n, orc.seed(20, 9), 9, 1000) if filtered else None
    want = orc.split(cols, 2, pred=pred, bound=499)
    s = Split(2, 1, n)
    s.run(cols, pred=pred, bound=499)
    assert s.cursors().tolist() == [len(want[0][0]), len(want[1][0])]
    for p in range(2):
        assert np.array_equal(canon(s.parent(p)), canon(want[p]))


@pytest.mark.parametrize("grid", [1, 3, 7])
def test_split_small_grids_cover_every_tile(gpu, orc, grid):
    n = 16384 * 18 + 5   # eighteen whole 16384-row tiles of one column and a ragged one
    cols = seeded_cols(orc, n, 1, 40)
    want = orc.split(cols, 2)
    s = Split(2, 1, n)
    s.run(cols, grid=grid)
    for p in range(2):
        assert np.array_equal(canon(s.parent(p)), canon(want[p]))


def test_split_overflow_is_counted_not_written(gpu, orc):
    n = 50000
    cols = seeded_cols(orc, n, 2, 50)
    want = orc.split(cols, 2)
    cap = 1000
    s = Split(2, 2, cap)
    s.run(cols)
    assert s.cursors().tolist() == [len(want[0][0]), len(want[1][0])]   # > cap: the overflow
    for p in range(2):
        got = canon(s.parent(p))
        assert got.shape[1] == cap
        # every written row is one of that parent's rows, none twice
        allowed = {tuple(r) for r in canon(want[p]).T.tolist()}
        rows = [tuple(r) for r in got.T.tolist()]
        assert len(set(rows)) == len(rows) and set(rows) <= allowed
        assert s.guard(p, 0) == -7 and s.guard(p, 1) == -7


def test_split_one_parent_keeps_even_rows_only(gpu, orc):
    x = np.arange(-50, 50, dtype=np.int32)
    s = Split(1, 1, 100)
    s.run([x])
    assert sorted(s.parent(0)[0].tolist()) == [v for v in range(-50, 50) if v % 2 == 0]


def test_split_large_vector_checksums(gpu, orc):
    """2^25 rows, filter at selectivity 1/2: cursors and per-parent sums
    equal numpy's, and every row of a parent has that parity."""
    n = 1 << 25
    x = orc.gen_bits(n, orc.seed(60, 0), 9, 31)
    v = orc.gen_bits(n, orc.seed(60, 1), 9, 16)
    pred = orc.gen_bits(n, orc.seed(60, 2), 9, 31)
    bound = (1 << 30) - 1
    s = Split(2, 2, n)
    s.run([x, v], pred, bound)
    keep = pred <= bound
    for p in range(2):
        sel = keep & ((x & 1) == p)
        got = s.parent(p)
        assert len(got[0]) == int(sel.sum())
        assert np.all((got[0] & 1) == p)
        assert int(got[0].astype(np.int64).sum()) == int(x[sel].astype(np.int64).sum())
        assert int(got[1].astype(np.int64).sum()) == int(v[sel].astype(np.int64).sum())


def test_split_rejects_bad_arguments(gpu):
    a = mgv.device_array(np.arange(16, dtype=np.int32))
    outs = [mgv.DeviceBuffer(64) for _ in range(4)]
    cnt = mgv.device_array(np.zeros(9, np.uint64))
    cp = (VP * 1)(a.ptr)
    op = (VP * 4)(*[o.ptr for o in outs])
    L = mgv.launch(0, 1024)
    for ncols, npar, cols, out in ((0, 2, cp, op), (5, 2, cp, op), (1, 0, cp, op), (1, 9, cp, op),
                                   (1, 2, None, op), (1, 2, cp, None), (1, 2, (VP * 1)(None), op),
                                   (1, 2, cp, (VP * 2)(outs[0].ptr, None))):
        with pytest.raises(mgv.MgvError):
            mgv.call("mgv_split", None, 0, cols, ncols, 16, npar, out, 16, cnt.vp, L, None)
    with pytest.raises(mgv.MgvError) as e:   # cursors not 8-byte aligned
        mgv.call("mgv_split", None, 0, cp, 1, 16, 2, op, 16, cnt.at(4), L, None)
    assert e.value.code == mgv.E_ALIGN
    mgv.sync()
    assert cnt.download(np.uint64, 9).tolist() == [0] * 9
