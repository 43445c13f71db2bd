"""GPU tests of the two-level partitioned probe with LDS-resident table slices
(mgv_hashjoin_probe_lds): COUNT and int64 SUM must equal the oracle's
chained-table join and the direct GPU probe, including duplicate build keys,
chains that run past a slice and wrap around the table, the key -1 (half of
the empty-slot pattern), a filter column, and a missing value column."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def build(keys, payload, lf):
    ht = mgv.HashTable(len(keys), lf)
    dk, dp = dev(keys), dev(payload)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, len(keys), 0,
             res.vp, mgv.launch(0, 256), None)
    mgv.sync()
    assert res.read().errors == 0
    return ht


def probe_lds(ht, rk, rv, pred=None, bound=0, fanout=64, grid=0, flags=0):
    n = len(rk)
    need = C.c_uint64()
    mgv.call("mgv_probe_lds_workspace_size", n, ht.nbuckets, fanout, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    dk = dev(rk)
    dv = dev(rv) if rv is not None else None
    dp = dev(pred) if pred is not None else None
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, dp.vp if dp else None, bound, dk.vp,
             dv.vp if dv else None, n, fanout, ws.vp, ws.nbytes, res.vp, mgv.Launch(grid, 1024, -1, flags), None)
    mgv.sync()
    r = res.read()
    return r.count, r.sum


def probe_direct(ht, rk, rv, pred=None, bound=0):
    dk = dev(rk)
    dv = dev(rv) if rv is not None else None
    dp = dev(pred) if pred is not None else None
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, dp.vp if dp else None, bound, dk.vp,
             dv.vp if dv else None, len(rk), 0, 0, res.vp, mgv.launch(0, 256), None)
    mgv.sync()
    r = res.read()
    return r.count, r.sum


@pytest.mark.parametrize("nb_keys,n,fanout,lf", [(1 << 16, 1 << 20, 64, 0.5), (1 << 20, 3 << 20, 64, 0.5),
                                                 (16384, 100003, 2, 0.5), (51200, 200000, 8, 0.75)])
def test_lds_probe_matches_oracle(gpu, orc, nb_keys, n, fanout, lf):
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 7)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 7, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 7, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 7, 16)
    ht = build(sk, sw, lf)
    got = probe_lds(ht, rk, rv, fanout=fanout)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    assert got == (want["count"], want["sum"]) == probe_direct(ht, rk, rv)


COUNTED = 1 << 12   # MGV_LAUNCH_COUNTED_PART


@pytest.mark.parametrize("nb_keys,n,fanout", [(1 << 16, 1 << 20, 64), (1 << 18, 3 << 20, 32), (50000, 123457, 4),
                                              (1 << 16, 1 << 20, 2048)])
def test_lds_probe_fill_and_counted_partitionings_agree(gpu, orc, nb_keys, n, fanout):
    """The default count-free partitioning (fixed-capacity regions filled by
    per-tile reservations) and the counted one give the oracle's result, with
    two levels and with one (fan-out 2048 on 2^13 buckets of 8 slots)."""
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 12)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 12, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 12, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 12, 16)
    ht = build(sk, sw, 0.5 if fanout != 2048 else 0.25)
    if fanout == 2048:
        assert ht.nbuckets >= 2048 and ht.nbuckets % 2048 == 0
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    got = probe_lds(ht, rk, rv, fanout=fanout)
    assert got == (want["count"], want["sum"]) == probe_lds(ht, rk, rv, fanout=fanout, flags=COUNTED)


@pytest.mark.parametrize("nb_keys,lf,fanout", [(1 << 22, 0.0625, 32), (1 << 21, 0.015625, 8)])
def test_lds_probe_more_slices_than_the_counted_partitioning_takes(gpu, orc, nb_keys, lf, fanout):
    """2^23 buckets (4096 slices of 2048 buckets: level 2 cuts each of 32
    regions into 128) and 2^24 buckets (8192 slices: 8 regions into 1024):
    the count-free partitioning takes them, the counted one (2048 slices at
    most) refuses them, and COUNT and SUM equal the oracle's and the direct
    probe's, with and without a filter."""
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 15)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 15, 16)
    n = 1 << 22
    rk = orc.gen_range(n, orc.seed(2, 2), 15, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 15, 16)
    c = orc.gen_bits(n, orc.seed(2, 4), 15, 31)
    ht = build(sk, sw, lf)
    assert ht.nbuckets in (1 << 23, 1 << 24)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    bound = (1 << 30) - 1
    sel = c <= bound
    want_f = orc.probe(t, rk[sel], rv[sel])
    orc.free(t)
    assert probe_lds(ht, rk, rv, fanout=fanout) == (want["count"], want["sum"]) == probe_direct(ht, rk, rv)
    assert probe_lds(ht, rk, rv, pred=c, bound=bound, fanout=fanout) == (want_f["count"], want_f["sum"])
    with pytest.raises(mgv.MgvError):
        probe_lds(ht, rk, rv, fanout=fanout, flags=COUNTED)


@pytest.mark.parametrize("kind", ["one_key", "zipf"])
@pytest.mark.parametrize("nb_keys,n", [(1 << 16, 1 << 20), (1 << 20, 1 << 21)])
def test_lds_probe_skewed_keys_overflow_their_regions(gpu, orc, kind, nb_keys, n):
    """Skewed probe keys overfill their partitions' regions: the excess rows
    go to the overflow areas, which the hashed global-table probe reads, and
    the result still equals the oracle's.  2^16 keys (2^14 buckets) take one
    level of partitions at both fan-outs; 2^20 keys (2^18 buckets) take two
    at fan-out 32 (level-1 overflow and level-2 overflow) and one at 256."""
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 13)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 13, 16)
    rv = orc.gen_bits(n, orc.seed(2, 3), 13, 16)
    if kind == "one_key":
        rk = orc.gen_range(n, orc.seed(2, 2), 13, 2 * nb_keys)
        rk[: (3 * n) // 5] = 12345          # 60 % of the rows on one key (one partition)
    else:
        T = orc.zipf_cdf(nb_keys, 1.2)
        rk = orc.gen_zipf(n, orc.seed(2, 2), 13, T)
    ht = build(sk, sw, 0.5)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    for fanout in (32, 256):
        assert probe_lds(ht, rk, rv, fanout=fanout) == (want["count"], want["sum"])
    assert probe_direct(ht, rk, rv) == (want["count"], want["sum"])


def test_lds_probe_with_filter_and_without_values(gpu, orc):
    nb_keys, n = 1 << 17, 1 << 21
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 8)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 8, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 8, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 8, 16)
    c = orc.gen_bits(n, orc.seed(3, 0), 8, 31)
    ht = build(sk, sw, 0.5)
    t = orc.table(sk, sw)
    for bound in ((1 << 30) - 1, 21474836, -1, 2**31 - 1):
        want = orc.probe(t, rk, rv, pred=c, bound=bound)
        assert probe_lds(ht, rk, rv, pred=c, bound=bound) == (want["count"], want["sum"])
    want = orc.probe(t, rk)
    orc.free(t)
    assert probe_lds(ht, rk, None) == (want["count"], want["sum"])


def test_lds_probe_duplicates_long_chains_and_wraparound(gpu, orc):
    """Load factor 0.9 with heavy duplicates: chains leave their slice (and the
    extra buckets loaded past it) and wrap from the last bucket to bucket 0."""
    nb_keys, n = 60000, 400000
    sk = orc.gen_range(nb_keys, orc.seed(2, 6), 0, 6000)          # ~10 copies of each key
    sw = orc.gen_bits(nb_keys, orc.seed(2, 7), 0, 16)
    rk = orc.gen_range(n, orc.seed(2, 8), 0, 7000)
    rv = orc.gen_bits(n, orc.seed(2, 9), 0, 16)
    ht = build(sk, sw, 0.9)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    for fanout in (2, 4):
        assert probe_lds(ht, rk, rv, fanout=fanout) == (want["count"], want["sum"]) == probe_direct(ht, rk, rv)


KEY_SLICES = 1 << 13   # MGV_LAUNCH_KEY_SLICES


@pytest.mark.parametrize("nb_keys,lf,key_range,n,fanout", [
    (1 << 20, 0.5, 0, 1 << 20, 32),             # 2^18 buckets: tags of 14 hash bits + 2 offset bits
    (1887436, 0.9, 200000, 1 << 20, 16),        # 2^18 buckets at load factor 0.9, ~9 copies per key
    (1 << 21, 0.5, 0, 1 << 21, 2),              # 2^19 buckets, offsets up to 6
    (1 << 23, 0.5, 3 << 20, 1 << 21, 64),       # 2^21 buckets, 2.7 copies per key
])
def test_tagged_slices_equal_key_slices_and_oracle(gpu, orc, nb_keys, lf, key_range, n, fanout):
    """Power-of-two tables of >= 2^18 buckets are probed through 16-bit slot
    tags (chain offset + low hash bits): the result equals the key slices'
    (MGV_LAUNCH_KEY_SLICES), the direct probe's and the oracle's, also at
    load factor 0.9 with duplicates, where chains run past the offsets a tag
    holds and continue in global memory."""
    if key_range:
        sk = orc.gen_range(nb_keys, orc.seed(2, 10), 14, key_range)
    else:
        sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 10), 14)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 11), 14, 16)
    rk = orc.gen_range(n, orc.seed(2, 12), 14, int(1.1 * (key_range or nb_keys)))
    rv = orc.gen_bits(n, orc.seed(2, 13), 14, 16)
    ht = build(sk, sw, lf)
    assert ht.nbuckets & (ht.nbuckets - 1) == 0 and ht.nbuckets >= 1 << 18
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    got = probe_lds(ht, rk, rv, fanout=fanout)
    assert got == (want["count"], want["sum"])
    assert probe_lds(ht, rk, rv, fanout=fanout, flags=KEY_SLICES) == got
    assert probe_direct(ht, rk, rv) == got
    ht.buf.close()


def test_lds_probe_key_minus_one(gpu, orc):
    nb_keys, n = 8192, 50000
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 9) - 1          # includes -1 (payload != -1)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 9, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 9, 2 * nb_keys) - 1
    rk[::97] = -1
    rv = orc.gen_bits(n, orc.seed(2, 3), 9, 16)
    ht = build(sk, sw, 0.5)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv)
    orc.free(t)
    assert probe_lds(ht, rk, rv, fanout=2) == (want["count"], want["sum"])


def test_lds_probe_few_blocks_and_rejects_unsplittable_tables(gpu, orc):
    nb_keys, n = 1 << 16, 1 << 19
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 10)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 10, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 10, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 10, 16)
    ht = build(sk, sw, 0.5)
    assert probe_lds(ht, rk, rv, fanout=16, grid=3) == probe_direct(ht, rk, rv)
    need = C.c_uint64()
    with pytest.raises(mgv.MgvError):
        mgv.call("mgv_probe_lds_workspace_size", 1000, 25, 2, C.byref(need))     # odd bucket count
    with pytest.raises(mgv.MgvError):   # 8192 slices: 4096 per level-1 region, above 1024
        mgv.call("mgv_probe_lds_workspace_size", 1000, 1 << 24, 2, C.byref(need))
    with pytest.raises(mgv.MgvError):   # 65536 slices, above 32768
        mgv.call("mgv_probe_lds_workspace_size", 1000, 1 << 27, 64, C.byref(need))
    mgv.call("mgv_probe_lds_workspace_size", 1000, 1 << 24, 64, C.byref(need))   # 8192 slices, 128 per region


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_scatter_shapes_give_identical_results(gpu, orc, shape):
    """Every scatter workgroup shape (MGV_LAUNCH_SCATTER) partitions the same
    rows: the LDS probe, the partitioned probe and the partitioned group-by
    give the oracle's results with each."""
    nb_keys, n = 1 << 16, (1 << 20) + 333
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 11)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 11, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 11, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 11, 16)
    c = orc.gen_bits(n, orc.seed(3, 0), 11, 31)
    bound = (1 << 30) - 1
    ht = build(sk, sw, 0.5)
    t = orc.table(sk, sw)
    want = orc.probe(t, rk, rv, pred=c, bound=bound)
    orc.free(t)
    flags = shape << 8
    # LDS probe
    need = C.c_uint64()
    mgv.call("mgv_probe_lds_workspace_size", n, ht.nbuckets, 16, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    dk, dv, dc = dev(rk), dev(rv), dev(c)
    res = mgv.ResultBuffer(2)
    mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, dc.vp, bound, dk.vp, dv.vp, n, 16, ws.vp, ws.nbytes,
             res.slot(0), mgv.Launch(0, 1024, -1, flags), None)
    # partitioned probe
    ws2 = mgv.partition_workspace(n, 128, 2)
    mgv.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, dc.vp, bound, dk.vp, dv.vp, n, 0, 0, 128,
             ws2.vp, ws2.nbytes, res.slot(1), mgv.Launch(0, 256, -1, flags), None)
    mgv.sync()
    for i in range(2):
        r = res.read(i)
        assert (r.count, r.sum) == (want["count"], want["sum"]), (i, shape)
    # partitioned group-by
    G = 20000
    gk = orc.gen_range(n, orc.seed(5, 0), 11, G)
    dg = dev(gk)
    g = mgv.DeviceBuffer(16 * G)
    mgv.call("mgv_fill_u64", g.vp, 2 * G, 0, None)
    mgv.call("mgv_groupby_workspace_size", n, G, C.byref(need))
    ws3 = mgv.DeviceBuffer(need.value)
    mgv.call("mgv_groupby_partitioned", dg.vp, dv.vp, n, g.vp, g.at(8 * G), G, ws3.vp, ws3.nbytes, None,
             mgv.Launch(0, 1024, -1, flags), None)
    mgv.sync()
    wc, wsum = orc.groupby(gk, rv, G)
    assert np.array_equal(g.download(np.uint64, G), wc) and np.array_equal(g.download(np.int64, G, 8 * G), wsum)


def _assert_prefix_filled(slots: np.ndarray, nbuckets: int, nkeys: int):
    """Every bucket's occupied slots (not the empty pattern (-1, -1)) are a
    prefix of its 8 slots, and every inserted row is in the table once."""
    occ = (slots != np.uint64(0xFFFFFFFFFFFFFFFF)).reshape(nbuckets, 8)
    c = occ.sum(axis=1)
    prefix = np.arange(8)[None, :] < c[:, None]
    assert np.array_equal(occ, prefix)
    assert int(c.sum()) == nkeys


@pytest.mark.parametrize("lf,dup", [(0.25, False), (0.5, False), (0.9, False), (0.9, True), (0.5, True)])
def test_build_fills_bucket_slots_in_order(gpu, orc, lf, dup):
    """The LDS probe stores a slice as keys plus a per-bucket fill count and
    treats slots [0, count) as the occupied ones: that holds only if the
    build claims a bucket's slots strictly in order.  Checked on tables built
    concurrently (the build kernel and the fused batch build), with heavy
    duplicates and chains that run over many buckets."""
    n = 200000
    keys = orc.gen_range(n, orc.seed(6, 0), 1, 5000) if dup else orc.gen_perm(n, n, orc.seed(6, 0), 1)
    pay = orc.gen_bits(n, orc.seed(6, 1), 1, 16)
    ht = build(keys, pay, lf)
    _assert_prefix_filled(ht.slots(), ht.nbuckets, n)
    # the same rows through the fused batch kernel's build
    ht2 = mgv.HashTable(n, lf)
    dk, dp = dev(keys), dev(pay)
    res = mgv.ResultBuffer(1)
    q = (mgv.Query * 1)()
    q[0] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, dk.ptr, dp.ptr, n, ht2.buf.ptr, ht2.nbuckets, 0, None, None,
                     res.ptr, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
    cnt = C.c_uint64()
    mgv.call("mgv_batch_plan", q, 1, 4096, None, 0, C.byref(cnt))
    items = (mgv.WorkItem * cnt.value)()
    mgv.call("mgv_batch_plan", q, 1, 4096, items, cnt.value, C.byref(cnt))
    dq = mgv.DeviceBuffer(C.sizeof(q)).upload(np.frombuffer(bytes(q), dtype=np.uint8))
    di = mgv.DeviceBuffer(C.sizeof(items)).upload(np.frombuffer(bytes(items), dtype=np.uint8))
    ctr = mgv.DeviceBuffer(256)
    mgv.call("mgv_fill_u32", ctr.vp, 64, 0, None)
    mgv.call("mgv_batch_run", dq.vp, di.vp, cnt.value, ctr.vp, mgv.launch(0, 256, 0), None)
    mgv.sync()
    assert res.read(0).errors == 0
    _assert_prefix_filled(ht2.slots(), ht2.nbuckets, n)


@pytest.mark.parametrize("n", [1, 3, 4, 255, 257, 4097, 9000])
def test_tiny_probes_through_both_partitioned_paths(gpu, orc, n):
    """Inputs smaller than one scatter tile, or not a multiple of 4 rows:
    the LDS probe and the one-level partitioned probe give the direct
    probe's COUNT and SUM, with and without a filter."""
    nb_keys = 1 << 14
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(2, 0), 12)
    sw = orc.gen_bits(nb_keys, orc.seed(2, 1), 12, 16)
    rk = orc.gen_range(n, orc.seed(2, 2), 12, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(2, 3), 12, 16)
    c = orc.gen_bits(n, orc.seed(3, 0), 12, 31)
    ht = build(sk, sw, 0.5)
    for pred, bound in ((None, 0), (c, (1 << 30) - 1)):
        want = probe_direct(ht, rk, rv, pred, bound)
        assert probe_lds(ht, rk, rv, pred=pred, bound=bound, fanout=4) == want
        ws = mgv.partition_workspace(n, 16, 2)
        dk, dv = dev(rk), dev(rv)
        dp = dev(pred) if pred is not None else None
        res = mgv.ResultBuffer()
        mgv.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, dp.vp if dp else None, bound, dk.vp,
                 dv.vp, n, 0, 0, 16, ws.vp, ws.nbytes, res.vp, mgv.launch(0, 256), None)
        mgv.sync()
        r = res.read()
        assert (r.count, r.sum) == want


def _gen(name, n, *args):
    buf = mgv.DeviceBuffer(4 * n)
    mgv.call(name, buf.vp, n, *args, None)
    return buf


@pytest.mark.parametrize("fanout", [32, 64])
def test_c2_table_geometry_lds_build_and_probe_equal_closed_form(gpu, orc, fanout):
    """C2's own table: 2^24 distinct keys (a seeded permutation) at load
    factor 0.5 is 2^22 buckets, so both LDS paths run with a level-2 fan-out
    of 2048 (the 2048-way count and scatter instances and 2048 LDS slices of
    2048 buckets).  2^25 probe rows of the C2 distribution; COUNT and SUM
    must equal the closed form COUNT = #{R.k < 2^24}, SUM = sum of R.v +
    w[perm^-1(R.k)], computed on the CPU by direct addressing."""
    ns, nr = 1 << 24, 1 << 25
    sk = _gen("mgv_gen_perm", ns, ns, mgv.seed(2, 0), 0)
    sw = _gen("mgv_gen_bits", ns, mgv.seed(2, 1), 0, 16)
    rk = _gen("mgv_gen_bits", nr, mgv.seed(2, 2), 0, 25)
    rv = _gen("mgv_gen_bits", nr, mgv.seed(2, 3), 0, 16)
    mgv.sync()
    hk, hw, hrk, hrv = sk.download(np.int32, ns), sw.download(np.int32, ns), rk.download(np.int32, nr), \
        rv.download(np.int32, nr)
    # the device generator is the oracle's (spot check of every column)
    assert np.array_equal(hk[:4096], orc.gen_perm(4096, ns, orc.seed(2, 0), 0))
    assert np.array_equal(hrk[-4096:], orc.gen_bits(nr, orc.seed(2, 2), 0, 25)[-4096:])
    assert np.array_equal(hw[:4096], orc.gen_bits(4096, orc.seed(2, 1), 0, 16))
    assert np.array_equal(hrv[:4096], orc.gen_bits(4096, orc.seed(2, 3), 0, 16))
    want = orc.join_dense(hk, hw, ns, hrk, hrv, threads=8)
    assert want[0] == int((hrk < ns).sum())

    ht = mgv.HashTable(ns, 0.5)
    assert ht.nbuckets == 1 << 22
    need = C.c_uint64()
    mgv.call("mgv_build_lds_workspace_size", ns, ht.nbuckets, fanout, C.byref(need))
    bws = mgv.DeviceBuffer(need.value)
    res = mgv.ResultBuffer(3)
    mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, sk.vp, sw.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, ns, 0,
             fanout, bws.vp, bws.nbytes, res.slot(0), mgv.Launch(0, 1024, -1, 2 << 8), None)
    mgv.sync()
    bws.close()
    assert res.read(0).errors == 0
    _assert_prefix_filled(ht.slots(), ht.nbuckets, ns)

    mgv.call("mgv_probe_lds_workspace_size", nr, ht.nbuckets, fanout, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    for rep in range(2):   # the workspace is reused as in the bench's steps
        res.clear(None)
        mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp, nr, fanout, ws.vp,
                 ws.nbytes, res.slot(1), mgv.Launch(0, 1024, -1, 2 << 8), None)
        mgv.sync()
        r = res.read(1)
        assert (r.count, r.sum) == want, (rep, fanout)
    mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp, nr, 0, 0, res.slot(2),
             mgv.launch(0, 256), None)
    mgv.sync()
    r = res.read(2)
    assert (r.count, r.sum) == want
