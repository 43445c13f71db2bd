"""GPU tests of mgv_hashjoin_probe_lds_noalloc, the probe through LDS table
slices that takes no workspace and allocates nothing (its partitionings use
a scratch array the library holds on each GPU).  Every call must give the
result of mgv_hashjoin_probe_lds (with a caller workspace) on the same
inputs, also when it follows a larger call that left its contents in the
scratch array, and a call too large for the array must fail before any
launch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20   # MGV_PROBE_NOALLOC_SCRATCH_BYTES


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def make_table(orc, nkeys, lf, stream):
    sk = orc.gen_perm(nkeys, nkeys, orc.seed(2, 0), stream)
    sw = orc.gen_bits(nkeys, orc.seed(2, 1), stream, 16)
    ht = mgv.HashTable(nkeys, lf)
    dk, dp = dev(sk), dev(sw)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, dk.vp, dp.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, nkeys, 0,
             res.vp, mgv.launch(0, 256), None)
    mgv.sync()
    assert res.read().errors == 0
    return ht, sk, sw


def with_workspace(ht, dk, dv, dp, bound, n, fanout):
    need = C.c_uint64()
    mgv.call("mgv_probe_lds_workspace_size", n, ht.nbuckets, fanout, C.byref(need))
    ws = mgv.DeviceBuffer(need.value)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, dp, bound, dk.vp, dv, n, fanout, ws.vp, ws.nbytes,
             res.vp, mgv.launch(0, 1024), None)
    mgv.sync()
    r = res.read()
    return need.value, (r.count, r.sum, r.errors)


def no_alloc(ht, dk, dv, dp, bound, n, fanout):
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe_lds_noalloc", ht.buf.vp, ht.nbuckets, dp, bound, dk.vp, dv, n, fanout, res.vp,
             mgv.launch(0, 1024), None)
    mgv.sync()
    r = res.read()
    return r.count, r.sum, r.errors


def test_no_alloc_probe_equals_workspace_probe_twice(gpu, orc):
    # first call: 2^20 probe rows against a 2^18-bucket table (two levels)
    ht, sk, sw = make_table(orc, 1 << 20, 0.5, 31)
    n1 = 1 << 20
    rk = orc.gen_range(n1, orc.seed(2, 2), 31, 2 << 20)
    rv = orc.gen_bits(n1, orc.seed(2, 3), 31, 16)
    dk, dv = dev(rk), dev(rv)
    need1, want1 = with_workspace(ht, dk, dv.vp, None, 0, n1, 32)
    assert need1 <= SCRATCH
    got1 = no_alloc(ht, dk, dv.vp, None, 0, n1, 32)
    assert got1 == want1
    t = orc.table(sk, sw)
    o = orc.probe(t, rk, rv)
    orc.free(t)
    assert got1[:2] == (o["count"], o["sum"])

    # second call, smaller and ragged, with a filter and no value column,
    # on another table: nothing may leak from the first call's scratch
    ht2, sk2, sw2 = make_table(orc, 40000, 0.75, 32)
    n2 = 100003
    rk2 = orc.gen_range(n2, orc.seed(2, 2), 32, 80000)
    c2 = orc.gen_bits(n2, orc.seed(3, 0), 32, 31)
    bound = (1 << 30) - 1
    dk2, dc2 = dev(rk2), dev(c2)
    _, want2 = with_workspace(ht2, dk2, None, dc2.vp, bound, n2, 4)
    got2 = no_alloc(ht2, dk2, None, dc2.vp, bound, n2, 4)
    assert got2 == want2
    t = orc.table(sk2, sw2)
    o2 = orc.probe(t, rk2, None, pred=c2, bound=bound)
    orc.free(t)
    assert got2[:2] == (o2["count"], o2["sum"])

    # and the first shape again after the second
    assert no_alloc(ht, dk, dv.vp, None, 0, n1, 32) == want1
    for b in (ht.buf, ht2.buf):
        b.close()


def test_no_alloc_probe_rejects_rows_beyond_the_scratch(gpu, orc):
    ht, _, _ = make_table(orc, 1 << 16, 0.5, 33)
    n = 8 << 20   # 16 bytes per row and level: more than the array holds
    need = C.c_uint64()
    mgv.call("mgv_probe_lds_workspace_size", n, ht.nbuckets, 32, C.byref(need))
    assert need.value > SCRATCH
    res = mgv.ResultBuffer()
    res.clear(None)
    mgv.sync()
    # the key pointer is never read: the call fails before any launch
    dummy = mgv.DeviceBuffer(256)
    rc = mgv.load().mgv_hashjoin_probe_lds_noalloc(ht.buf.vp, ht.nbuckets, None, 0, dummy.vp, None, n, 32, res.vp,
                                                   mgv.launch(0, 1024), None)
    assert rc == mgv.E_CAPACITY
    assert b"scratch" in mgv.load().mgv_last_error_string()
    mgv.sync()
    r = res.read()
    assert (r.count, r.sum) == (0, 0)
    ht.buf.close()
