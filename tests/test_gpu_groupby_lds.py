"""GPU tests of the probe + group-by through LDS table slices
(mgv_hashjoin_probe_groupby_lds): per-group COUNT and int64 SUM must equal
the oracle's chained-table probe + group-by and the direct GPU kernel
(mgv_hashjoin_probe_groupby), with a filter, skewed (Zipf) probe keys whose
hot partitions overflow their regions, duplicate build keys, out-of-range
group payloads, and up to 4096 groups."""
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


def groups(ngroups):
    g = mgv.DeviceBuffer(16 * ngroups)
    mgv.call("mgv_fill_u64", g.vp, 2 * ngroups, 0, None)
    return g


def read(g, ngroups):
    mgv.sync()
    return g.download(np.uint64, ngroups), g.download(np.int64, ngroups, 8 * ngroups)


def probe_groupby_lds(ht, rk, rv, pred, bound, ngroups, fanout=32, flags=2 << 8):
    n = len(rk)
    need = C.c_uint64()
    mgv.call("mgv_probe_groupby_lds_workspace_size", n, ht.nbuckets, fanout, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    dk, dv = dev(rk), dev(rv)
    dp = dev(pred) if pred is not None else None
    g = groups(ngroups)
    mgv.call("mgv_hashjoin_probe_groupby_lds", ht.buf.vp, ht.nbuckets, dp.vp if dp else None, bound, dk.vp, dv.vp, n,
             g.vp, g.at(8 * ngroups), ngroups, fanout, ws.vp, ws.nbytes, mgv.Launch(0, 1024, -1, flags), None)
    return read(g, ngroups)


def probe_groupby_direct(ht, rk, rv, pred, bound, ngroups):
    dk, dv, dp = dev(rk), dev(rv), dev(pred)
    g = groups(ngroups)
    mgv.call("mgv_hashjoin_probe_groupby", ht.buf.vp, ht.nbuckets, dp.vp, bound, dk.vp, dv.vp, len(rk), g.vp,
             g.at(8 * ngroups), ngroups, mgv.launch(0, 256), None)
    return read(g, ngroups)


@pytest.mark.parametrize("nb_keys,n,ngroups,zipf,fanout", [
    (1 << 18, 1 << 21, 1024, 1.0, 32),      # the skewed join shape, scaled down
    (1 << 18, 1 << 21, 1024, 0.0, 64),
    (1 << 16, 700001, 4096, 1.0, 16),
    (1 << 20, 1 << 22, 1024, 1.0, 2048),    # one level of 2048 partitions
])
def test_lds_probe_groupby_matches_oracle(gpu, orc, nb_keys, n, ngroups, zipf, fanout):
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(3, 0), 21)
    sg = orc.gen_range(nb_keys, orc.seed(3, 1), 21, ngroups)       # S.g: the group of each build key
    if zipf > 0:
        T = orc.zipf_cdf(nb_keys, zipf)
        rk = orc.gen_zipf(n, orc.seed(3, 2), 21, T, orc.seed(3, 3), 21)   # Zipf over the build keys
    else:
        rk = orc.gen_range(n, orc.seed(3, 2), 21, 2 * nb_keys)
    rv = orc.gen_bits(n, orc.seed(3, 4), 21, 16)
    c = orc.gen_bits(n, orc.seed(3, 5), 21, 31)
    ht = build(sk, sg, 0.5)
    t = orc.table(sk, sg)
    for bound in ((1 << 30) - 1, 2**31 - 1):
        wc, ws = orc.probe_groupby(t, rk, rv, c, bound, ngroups)
        gc, gs = probe_groupby_lds(ht, rk, rv, c, bound, ngroups, fanout)
        assert np.array_equal(gc, wc) and np.array_equal(gs, ws)
        if ht.nbuckets >= 1 << 18:   # tagged slices by default: the key slices agree
            kc, ks = probe_groupby_lds(ht, rk, rv, c, bound, ngroups, fanout, flags=(2 << 8) | (1 << 13))
            assert np.array_equal(kc, wc) and np.array_equal(ks, ws)
        dc, ds = probe_groupby_direct(ht, rk, rv, c, bound, ngroups)
        assert np.array_equal(dc, wc) and np.array_equal(ds, ws)
    orc.free(t)


def test_lds_probe_groupby_duplicates_and_out_of_range_groups(gpu, orc):
    """Duplicate build keys (several matches per row, in different groups)
    at load factor 0.75, payloads past ngroups dropped, no filter."""
    nb_keys, n, ngroups = 200000, 1 << 20, 1000
    sk = orc.gen_range(nb_keys, orc.seed(3, 6), 0, 50000)             # ~4 copies of each key
    sg = orc.gen_range(nb_keys, orc.seed(3, 7), 0, ngroups + 100)     # 10 % of payloads out of range
    rk = orc.gen_range(n, orc.seed(3, 8), 0, 60000)
    rv = orc.gen_bits(n, orc.seed(3, 9), 0, 31) - (1 << 30)           # signed values
    ht = build(sk, sg, 0.75)
    t = orc.table(sk, sg)
    wc, ws = orc.probe_groupby(t, rk, rv, None, 0, ngroups)
    orc.free(t)
    gc, gs = probe_groupby_lds(ht, rk, rv, None, 0, ngroups, fanout=8)
    assert np.array_equal(gc, wc) and np.array_equal(gs, ws)


def test_lds_probe_groupby_rejects_too_many_groups(gpu):
    ht = mgv.HashTable(1 << 12, 0.5)
    x = dev(np.zeros(1024, np.int32))
    g = groups(8192)
    with pytest.raises(mgv.MgvError) as e:
        mgv.call("mgv_hashjoin_probe_groupby_lds", ht.buf.vp, ht.nbuckets, None, 0, x.vp, x.vp, 1024, g.vp,
                 g.at(8 * 8192), 8192, 2, None, 0, mgv.Launch(0, 1024, -1, 0), None)
    assert e.value.code == mgv.E_INVALID


@pytest.mark.parametrize("nb_keys", [1 << 16, 1 << 20])
def test_lds_probe_groupby_one_hot_key(gpu, orc, nb_keys):
    """60 % of the probe rows on one key: its partition's region and the
    overflow area are full of that key, so whole waves add to one group (the
    wave-combined path of the group accumulator) -- in the LDS probe and in
    the hashed probe of the overflow rows; one level (2^14 buckets) and two
    (2^18 buckets).  Equal to the oracle and the direct kernel."""
    n, ngroups = 1 << 21, 1024
    sk = orc.gen_perm(nb_keys, nb_keys, orc.seed(3, 10), 22)
    sg = orc.gen_range(nb_keys, orc.seed(3, 11), 22, ngroups)
    rk = orc.gen_range(n, orc.seed(3, 12), 22, 2 * nb_keys)
    rk[: (3 * n) // 5] = 4321
    rv = orc.gen_bits(n, orc.seed(3, 13), 22, 16)
    c = orc.gen_bits(n, orc.seed(3, 14), 22, 31)
    ht = build(sk, sg, 0.5)
    t = orc.table(sk, sg)
    bound = (3 << 29) - 1     # 75 %
    wc, ws = orc.probe_groupby(t, rk, rv, c, bound, ngroups)
    orc.free(t)
    for fanout in (32, 256):
        gc, gs = probe_groupby_lds(ht, rk, rv, c, bound, ngroups, fanout)
        assert np.array_equal(gc, wc) and np.array_equal(gs, ws)
    dc, ds = probe_groupby_direct(ht, rk, rv, c, bound, ngroups)
    assert np.array_equal(dc, wc) and np.array_equal(ds, ws)
