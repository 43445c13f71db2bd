"""GPU tests of the build through LDS table slices (mgv_hashjoin_build_lds):
the table must hold the same {key, payload} multiset as the direct build's,
keep every bucket's occupied slots a prefix (what the LDS probe relies on),
and answer probes exactly as the oracle's chained table does -- with unique
and duplicate keys, chains that leave a slice or wrap around the table end
(the overflow pass), a filter, no payload column, row-id payloads, tiny
inputs, rejected rows and a build on top of an existing table."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def build_direct(ht, keys, pay, mode=mgv.MGV_PAYLOAD_COLUMN, pred=None, bound=0):
    dk = dev(keys)
    dp = dev(pay) if pay is not None else None
    dc = dev(pred) if pred is not None else None
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp if dp else None, mode,
             dc.vp if dc else None, bound, len(keys), 0, res.vp, mgv.launch(0, 256), None)
    mgv.sync()
    return res.read().errors


COUNTED = 1 << 12   # MGV_LAUNCH_COUNTED_PART: counted partitionings instead of the count-free ones


def build_lds(ht, keys, pay, mode=mgv.MGV_PAYLOAD_COLUMN, pred=None, bound=0, fanout=4, shape=2, flags=0):
    n = len(keys)
    need = C.c_uint64()
    mgv.call("mgv_build_lds_workspace_size", n, ht.nbuckets, fanout, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    dk = dev(keys)
    dp = dev(pay) if pay is not None else None
    dc = dev(pred) if pred is not None else None
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp if dp else None, mode,
             dc.vp if dc else None, bound, n, 0, fanout, ws.vp, ws.nbytes, res.vp,
             mgv.Launch(0, 1024, -1, (shape << 8) | flags), None)
    mgv.sync()
    return res.read().errors


def occupied(ht):
    s = ht.slots()
    return np.sort(s[s != EMPTY])


def assert_prefix_filled(ht):
    occ = (ht.slots() != EMPTY).reshape(ht.nbuckets, 8)
    c = occ.sum(axis=1)
    assert np.array_equal(occ, np.arange(8)[None, :] < c[:, None])


def probe(ht, rk, rv):
    dk, dv = dev(rk), dev(rv)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, None, 0, dk.vp, dv.vp, len(rk), 0, 0, res.vp,
             mgv.launch(0, 256), None)
    mgv.sync()
    r = res.read()
    return r.count, r.sum


@pytest.mark.parametrize("flags", [0, COUNTED], ids=["fill", "counted"])
@pytest.mark.parametrize("n,lf,dup,fanout", [(1 << 18, 0.5, False, 4), (1 << 20, 0.25, False, 16),
                                             (300000, 0.9, False, 2), (300000, 0.9, True, 2),
                                             (120000, 0.75, True, 8)])
def test_lds_build_matches_direct_build_and_oracle(gpu, orc, n, lf, dup, fanout, flags):
    keys = orc.gen_range(n, orc.seed(7, 0), 3, n // 40) if dup else orc.gen_perm(n, n, orc.seed(7, 0), 3)
    pay = orc.gen_bits(n, orc.seed(7, 1), 3, 16)
    a, b = mgv.HashTable(n, lf), mgv.HashTable(n, lf)
    assert build_direct(a, keys, pay) == 0
    assert build_lds(b, keys, pay, fanout=fanout, flags=flags) == 0
    assert np.array_equal(occupied(a), occupied(b))
    assert_prefix_filled(b)
    rk = orc.gen_range(4 * n, orc.seed(7, 2), 3, 2 * max(1, int(keys.max()) + 1))
    rv = orc.gen_bits(4 * n, orc.seed(7, 3), 3, 16)
    t = orc.table(keys, pay)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    assert probe(b, rk, rv) == probe(a, rk, rv) == (want["count"], want["sum"])


def test_lds_build_more_slices_than_the_counted_partitioning_takes(gpu, orc):
    """2^23 buckets (4096 slices of 2048) and 2^24 buckets (8192 slices):
    beyond the counted partitioning's 2048 slices, so the count-free one
    builds them (level 2 cutting each level-1 region into 128, then 1024
    slices); the table equals the direct build's and answers probes as the
    direct build's does.  Low load factors keep the key columns small."""
    for nkeys, lf, F1 in (((1 << 22), 0.0625, 32), ((1 << 21), 0.015625, 8)):
        keys = orc.gen_perm(nkeys, nkeys, orc.seed(7, 9), 9)
        pay = orc.gen_bits(nkeys, orc.seed(7, 10), 9, 16)
        a, b = mgv.HashTable(nkeys, lf), mgv.HashTable(nkeys, lf)
        assert b.nbuckets in (1 << 23, 1 << 24)
        with pytest.raises(mgv.MgvError):   # the counted partitioning stops at 2048 slices
            build_lds(b, keys, pay, fanout=F1, flags=COUNTED)
        assert build_direct(a, keys, pay) == 0
        b.clear()
        assert build_lds(b, keys, pay, fanout=F1) == 0
        assert np.array_equal(occupied(a), occupied(b))
        assert_prefix_filled(b)
        rk = orc.gen_range(1 << 22, orc.seed(7, 11), 9, 2 * nkeys)
        rv = orc.gen_bits(1 << 22, orc.seed(7, 12), 9, 16)
        assert probe(b, rk, rv) == probe(a, rk, rv)
        del a, b


def test_lds_build_chains_wrap_around_the_table_end(gpu, orc):
    """Keys whose home buckets are the table's last ones, duplicated so their
    chains run off the last slice and wrap to bucket 0."""
    nb_keys = 1 << 14
    ht = mgv.HashTable(nb_keys, 0.9)
    pool = np.arange(1 << 16, dtype=np.uint32)
    h = pool.copy()                       # fmix32, only to pick keys near the end
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x85EBCA6B)
    h ^= h >> np.uint32(13)
    h *= np.uint32(0xC2B2AE35)
    h ^= h >> np.uint32(16)
    bucket = (h.astype(np.uint64) * np.uint64(ht.nbuckets)) >> np.uint64(32)
    tail_keys = pool[bucket >= ht.nbuckets - 4][:64].astype(np.int64)
    keys = np.repeat(tail_keys, 20)[:nb_keys].astype(np.int32)
    pay = orc.gen_bits(len(keys), orc.seed(7, 5), 4, 16)
    a, b = mgv.HashTable(nb_keys, 0.9), mgv.HashTable(nb_keys, 0.9)
    assert build_direct(a, keys, pay) == 0
    assert build_lds(b, keys, pay, fanout=2) == 0
    assert np.array_equal(occupied(a), occupied(b))
    assert_prefix_filled(b)
    occ0 = (b.slots() != EMPTY).reshape(b.nbuckets, 8)[0].sum()
    assert occ0 > 0   # the chains did wrap
    rk = np.concatenate([tail_keys, tail_keys + 1]).astype(np.int32)
    rv = np.arange(len(rk), dtype=np.int32)
    assert probe(b, rk, rv) == probe(a, rk, rv)


def test_lds_build_filter_no_payload_rowid_and_rejects(gpu, orc):
    n = 100000
    keys = orc.gen_perm(n, n, orc.seed(7, 0), 5) - 1          # includes key -1
    pay = orc.gen_bits(n, orc.seed(7, 1), 5, 16)
    pay[keys == -1] = -1                                      # the empty pattern: rejected
    c = orc.gen_bits(n, orc.seed(7, 2), 5, 31)
    bound = (1 << 30) - 1
    for mode, p, pred in ((mgv.MGV_PAYLOAD_COLUMN, pay, c), (mgv.MGV_PAYLOAD_NONE, None, None),
                          (mgv.MGV_PAYLOAD_ROWID, None, c), (mgv.MGV_PAYLOAD_COLUMN, pay, None)):
        a, b = mgv.HashTable(n, 0.5), mgv.HashTable(n, 0.5)
        ea = build_direct(a, keys, p, mode, pred, bound)
        eb = build_lds(b, keys, p, mode, pred, bound)
        assert ea == eb
        assert np.array_equal(occupied(a), occupied(b))
        assert_prefix_filled(b)
    # the rejected row is counted only when it passes the filter (or there is none)
    assert eb == 1


@pytest.mark.parametrize("n", [1, 5, 255, 4099])
def test_lds_build_tiny(gpu, orc, n):
    keys = orc.gen_perm(n, n, orc.seed(7, 0), 6)
    pay = orc.gen_bits(n, orc.seed(7, 1), 6, 16)
    a, b = mgv.HashTable(4096, 0.5), mgv.HashTable(4096, 0.5)   # 1024 buckets: room for 8192 rows
    build_direct(a, keys, pay)
    build_lds(b, keys, pay, fanout=2)
    assert np.array_equal(occupied(a), occupied(b))


def test_lds_build_on_top_of_an_existing_table(gpu, orc):
    """Half the rows built directly, the other half through LDS slices into
    the same table: the same table contents as one build of all rows."""
    n = 200000
    keys = orc.gen_range(n, orc.seed(7, 0), 7, n // 4)
    pay = orc.gen_bits(n, orc.seed(7, 1), 7, 16)
    a, b = mgv.HashTable(n, 0.75), mgv.HashTable(n, 0.75)
    build_direct(a, keys, pay)
    build_direct(b, keys[: n // 2], pay[: n // 2])
    build_lds(b, keys[n // 2:], pay[n // 2:], fanout=4)
    assert np.array_equal(occupied(a), occupied(b))
    assert_prefix_filled(b)
    rk = orc.gen_range(n, orc.seed(7, 2), 7, n // 2)
    rv = orc.gen_bits(n, orc.seed(7, 3), 7, 16)
    assert probe(a, rk, rv) == probe(b, rk, rv)


def test_lds_build_rejects_bad_arguments(gpu, orc):
    n = 8192   # 2048 buckets at load factor 0.5: splits into slices
    keys = orc.gen_perm(n, n, orc.seed(7, 0), 8)
    pay = orc.gen_bits(n, orc.seed(7, 1), 8, 16)
    ht = mgv.HashTable(n, 0.5)
    dk, dp = dev(keys), dev(pay)
    need = C.c_uint64()
    mgv.call("mgv_build_lds_workspace_size", n, ht.nbuckets, 4, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    L = mgv.Launch(0, 1024, -1, 2 << 8)
    with pytest.raises(mgv.MgvError):   # workspace one byte short
        mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, n, 0,
                 4, ws.vp, need.value - 1, None, L, None)
    with pytest.raises(mgv.MgvError):   # unknown payload mode
        mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp, 7, None, 0, n, 0, 4, ws.vp,
                 ws.nbytes, None, L, None)
    with pytest.raises(mgv.MgvError):   # payload column missing
        mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, dk.vp, None, mgv.MGV_PAYLOAD_COLUMN, None, 0, n, 0,
                 4, ws.vp, ws.nbytes, None, L, None)
    with pytest.raises(mgv.MgvError):   # bucket count that does not split into slices
        mgv.call("mgv_build_lds_workspace_size", n, 4099, 4, C.byref(need))
    with pytest.raises(mgv.MgvError):   # more rows than slots
        mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, 16, dk.vp, dp.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, n, 0, 4,
                 ws.vp, ws.nbytes, None, L, None)
    mgv.sync()
    # the table was not touched by the rejected calls
    assert not np.any(ht.slots() != EMPTY)
