"""GPU tests of the fused multi-query batch kernel, the device buffer pool /
lock-free stack, and the device->host result channel, through the C ABI."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

INT_MIN, INT_MAX = -2**31, 2**31 - 1


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def make_queries(orc, nq, seed_cfg=7, max_log=15):
    """Small C4-shaped query set: select (c <= t) -> join (k) -> sum(v + w).
    Returns host columns, device buffers and per-query oracle results."""
    rng = np.random.default_rng(seed_cfg)
    qs = []
    for q in range(nq):
        P = int(2 ** rng.uniform(8, max_log)) + int(rng.integers(0, 7))
        B = max(16, P // 16)
        sk = orc.gen_perm(B, B, orc.seed(4, 3), q)
        sw = orc.gen_bits(B, orc.seed(4, 4), q, 16)
        rk = orc.gen_range(P, orc.seed(4, 1), q, 2 * B)
        rv = orc.gen_bits(P, orc.seed(4, 2), q, 16)
        c = orc.gen_bits(P, orc.seed(4, 0), q, 31)
        bound = int(rng.integers(INT_MIN // 2, INT_MAX))
        t = orc.table(sk, sw)
        want = orc.probe(t, rk, rv, pred=c, bound=bound)
        orc.free(t)
        qs.append(dict(P=P, B=B, sk=sk, sw=sw, rk=rk, rv=rv, c=c, bound=bound, want=want,
                       d={k: dev(v) for k, v in dict(sk=sk, sw=sw, rk=rk, rv=rv, c=c).items()}))
    return qs


def run_batch(gpu, qs, vector_rows=1024, compact=1, block=256):
    nq = len(qs)
    tables = [mgv.HashTable(q["B"], 0.5) for q in qs]
    res = mgv.ResultBuffer(2 * nq)   # [0, nq): builds, [nq, 2nq): probes
    build_q = (mgv.Query * nq)()
    probe_q = (mgv.Query * nq)()
    for i, (q, t) in enumerate(zip(qs, tables)):
        d = q["d"]
        build_q[i] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, d["sk"].ptr, d["sw"].ptr, q["B"], t.buf.ptr, t.nbuckets, 0,
                               None, None, res.ptr + 64 * i, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
        probe_q[i] = mgv.Query(mgv.MGV_Q_PROBE_SUM, q["bound"], d["c"].ptr, d["rk"].ptr, d["rv"].ptr, q["P"],
                               t.buf.ptr, t.nbuckets, 0, None, None, res.ptr + 64 * (nq + i), 0, 0, 0)
    out = {}
    for name, qarr in (("build", build_q), ("probe", probe_q)):
        n = C.c_uint64()
        gpu.call("mgv_batch_plan", qarr, nq, vector_rows, None, 0, C.byref(n))
        items = (mgv.WorkItem * n.value)()
        gpu.call("mgv_batch_plan", qarr, nq, vector_rows, items, n.value, C.byref(n))
        dq = mgv.DeviceBuffer(C.sizeof(qarr)).upload(np.frombuffer(bytes(qarr), dtype=np.uint8))
        di = mgv.DeviceBuffer(max(1, C.sizeof(items))).upload(np.frombuffer(bytes(items), dtype=np.uint8))
        out[name] = (dq, di, n.value)
    counter = mgv.DeviceBuffer(64)
    L = mgv.launch(0, block, compact)
    gpu.call("mgv_batch_run", out["build"][0].vp, out["build"][1].vp, out["build"][2], counter.vp, L, None)
    gpu.call("mgv_batch_run", out["probe"][0].vp, out["probe"][1].vp, out["probe"][2], counter.vp, L, None)
    mgv.sync()
    return [res.read(nq + i) for i in range(nq)], [res.read(i) for i in range(nq)], tables


@pytest.mark.parametrize("vector_rows,compact,block", [(256, 1, 256), (1024, 0, 64), (4096, 1, 1024), (2048, 1, 128),
                                                      (4096, 0, 512)])
def test_fused_batch_matches_oracle(gpu, orc, vector_rows, compact, block):
    qs = make_queries(orc, 24)
    probes, builds, _ = run_batch(gpu, qs, vector_rows, compact, block)
    for q, r, b in zip(qs, probes, builds):
        assert b.errors == 0
        assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), (q["P"], q["bound"])


def test_fused_batch_each_query_alone_equals_in_mix(gpu, orc):
    qs = make_queries(orc, 12, seed_cfg=11)
    mix, _, _ = run_batch(gpu, qs)
    for i, q in enumerate(qs):
        alone, _, _ = run_batch(gpu, [q])
        assert (alone[0].count, alone[0].sum) == (mix[i].count, mix[i].sum) == (q["want"]["count"],
                                                                                q["want"]["sum"])


def test_fused_batch_mixed_kinds(gpu, orc):
    """SELECT_SUM, GROUPBY and PROBE_GROUPBY queries side by side in one launch."""
    n = 50001
    a = orc.gen_bits(n, orc.seed(1, 0), 0, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 0, 31)
    G = 1 << 12
    T = orc.zipf_cdf(G, 0.8)
    k = orc.gen_zipf(n, orc.seed(5, 0), 0, T)
    v = orc.gen_bits(n, orc.seed(5, 1), 0, 16)
    sk = orc.gen_perm(G, G, orc.seed(3, 3), 0)
    sg = orc.gen_bits(G, orc.seed(3, 4), 0, 6)
    da, db, dk, dv, dsk, dsg = dev(a), dev(b), dev(k), dev(v), dev(sk), dev(sg)
    ht = mgv.HashTable(G, 0.5)
    res = mgv.ResultBuffer(4)
    gpu.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, dsk.vp, dsg.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, G, 0,
             res.slot(3), mgv.launch(), None)
    gc1, gs1 = mgv.DeviceBuffer(8 * G), mgv.DeviceBuffer(8 * G)
    gc2, gs2 = mgv.DeviceBuffer(8 * 64), mgv.DeviceBuffer(8 * 64)
    for bb, m in ((gc1, G), (gs1, G), (gc2, 64), (gs2, 64)):
        gpu.call("mgv_fill_u64", bb.vp, m, 0, None)
    bound = (1 << 30) - 1
    qarr = (mgv.Query * 3)()
    qarr[0] = mgv.Query(mgv.MGV_Q_SELECT_SUM, bound, db.ptr, None, da.ptr, n, None, 0, 0, None, None, res.ptr, 0, 0, 0)
    qarr[1] = mgv.Query(mgv.MGV_Q_GROUPBY, 0, None, dk.ptr, dv.ptr, n, None, 0, G, gc1.ptr, gs1.ptr, res.ptr + 64,
                        0, 0, 0)
    qarr[2] = mgv.Query(mgv.MGV_Q_PROBE_GROUPBY, bound, db.ptr, dk.ptr, dv.ptr, n, ht.buf.ptr, ht.nbuckets, 64,
                        gc2.ptr, gs2.ptr, res.ptr + 128, 0, 0, 0)
    cnt = C.c_uint64()
    gpu.call("mgv_batch_plan", qarr, 3, 512, None, 0, C.byref(cnt))
    items = (mgv.WorkItem * cnt.value)()
    gpu.call("mgv_batch_plan", qarr, 3, 512, items, cnt.value, C.byref(cnt))
    dq = mgv.DeviceBuffer(C.sizeof(qarr)).upload(np.frombuffer(bytes(qarr), dtype=np.uint8))
    di = mgv.DeviceBuffer(C.sizeof(items)).upload(np.frombuffer(bytes(items), dtype=np.uint8))
    counter = mgv.DeviceBuffer(64)
    gpu.call("mgv_batch_run", dq.vp, di.vp, cnt.value, counter.vp, mgv.launch(0, 256, 1), None)
    mgv.sync()
    c0, s0, _ = orc.select_sum(b, a, bound)
    r0 = res.read(0)
    assert (r0.count, r0.sum) == (c0, s0)
    wc, ws = orc.groupby(k, v, G)
    assert np.array_equal(gc1.download(np.uint64, G), wc) and np.array_equal(gs1.download(np.int64, G), ws)
    t = orc.table(sk, sg)
    wc2, ws2 = orc.probe_groupby(t, k, v, b, bound, 64)
    orc.free(t)
    assert np.array_equal(gc2.download(np.uint64, 64), wc2) and np.array_equal(gs2.download(np.int64, 64), ws2)


def test_batch_plan_validation(gpu):
    q = (mgv.Query * 1)()
    q[0] = mgv.Query(99, 0, None, None, None, 10, None, 0, 0, None, None, None, 0, 0, 0)
    n = C.c_uint64()
    with pytest.raises(mgv.MgvError) as e:
        gpu.call("mgv_batch_plan", q, 1, 512, None, 0, C.byref(n))
    assert e.value.code == mgv.E_INVALID
    with pytest.raises(mgv.MgvError):
        gpu.call("mgv_batch_plan", q, 1, 100, None, 0, C.byref(n))   # not a multiple of 256


def test_fused_batch_graph_replays(gpu, orc):
    """The build + probe launches captured in one hipGraph give the same
    results on every replay (tables and results re-initialised by fill
    kernels inside the graph)."""
    qs = make_queries(orc, 8, seed_cfg=5, max_log=13)
    probes, _, tables = run_batch(gpu, qs)
    nq = len(qs)
    res = mgv.ResultBuffer(2 * nq)
    bq, pq = (mgv.Query * nq)(), (mgv.Query * nq)()
    for i, (q, t) in enumerate(zip(qs, tables)):
        d = q["d"]
        bq[i] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, d["sk"].ptr, d["sw"].ptr, q["B"], t.buf.ptr, t.nbuckets, 0, None,
                          None, res.ptr + 64 * i, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
        pq[i] = mgv.Query(mgv.MGV_Q_PROBE_SUM, q["bound"], d["c"].ptr, d["rk"].ptr, d["rv"].ptr, q["P"], t.buf.ptr,
                          t.nbuckets, 0, None, None, res.ptr + 64 * (nq + i), 0, 0, 0)
    bufs = []
    for qarr in (bq, pq):
        n = C.c_uint64()
        gpu.call("mgv_batch_plan", qarr, nq, 1024, None, 0, C.byref(n))
        items = (mgv.WorkItem * n.value)()
        gpu.call("mgv_batch_plan", qarr, nq, 1024, items, n.value, C.byref(n))
        bufs.append((mgv.DeviceBuffer(C.sizeof(qarr)).upload(np.frombuffer(bytes(qarr), dtype=np.uint8)),
                     mgv.DeviceBuffer(C.sizeof(items)).upload(np.frombuffer(bytes(items), dtype=np.uint8)),
                     n.value))
    st = mgv.Stream()
    counter = mgv.DeviceBuffer(64)
    gpu.call("mgv_graph_begin_capture", st.vp)
    for t in tables:
        t.clear(st.vp)
    res.clear(st.vp)
    for dq, di, n in bufs:
        gpu.call("mgv_batch_run", dq.vp, di.vp, n, counter.vp, mgv.launch(0, 256, 1), st.vp)
    g = C.c_void_p()
    gpu.call("mgv_graph_end_capture", st.vp, C.byref(g))
    for rep in range(4):
        gpu.call("mgv_graph_launch", g, st.vp)
        st.sync()
        for i, q in enumerate(qs):
            r = res.read(nq + i)
            assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), rep
        # eager work between replays
        res.clear(st.vp)
        st.sync()
    gpu.call("mgv_graph_destroy", g)
    st.close()


# ----------------------------------------------------------------- pool

def test_pool_acquire_release_and_end_of_stream(gpu):
    pool = C.c_void_p()
    gpu.call("mgv_pool_create", 3, 1024, C.byref(pool))
    b = C.c_int32()
    got = []
    gpu.call("mgv_pool_register_producer", pool, None)
    for _ in range(3):
        gpu.call("mgv_pool_acquire", pool, C.byref(b), None)
        got.append(b.value)
    assert sorted(got) == [0, 1, 2]
    # empty pool with a producer registered: retries run out
    gpu.call("mgv_pool_acquire", pool, C.byref(b), None)
    assert b.value == mgv.POOL_RETRY
    # no producer and nothing available: end of stream
    gpu.call("mgv_pool_unregister_producer", pool, None)
    gpu.call("mgv_pool_acquire", pool, C.byref(b), None)
    assert b.value == mgv.POOL_END
    for x in got:
        gpu.call("mgv_pool_release", pool, x, None)
    av, pr, stk = C.c_uint32(), C.c_uint32(), C.c_uint32()
    gpu.call("mgv_pool_stats", pool, C.byref(av), C.byref(pr), C.byref(stk), None)
    assert (av.value, pr.value, stk.value) == (3, 0, 3)
    with pytest.raises(mgv.MgvError):
        gpu.call("mgv_pool_release", pool, 7, None)
    gpu.call("mgv_pool_destroy", pool)


def test_pool_stack_concurrent_stress(gpu):
    pool = C.c_void_p()
    nbuf = 64
    gpu.call("mgv_pool_create", nbuf, 256, C.byref(pool))
    gpu.call("mgv_pool_register_producer", pool, None)
    ops = mgv.DeviceBuffer(32)
    gpu.call("mgv_fill_u64", ops.vp, 4, 0, None)
    blocks, iters = 512, 200
    gpu.call("mgv_pool_stress", pool, blocks, iters, ops.vp, None)
    mgv.sync()
    ok, bad, retry, end = ops.download(np.uint64, 4).tolist()
    assert bad == 0 and end == 0
    assert ok + retry == blocks * 4 * iters and ok > 0
    av, pr, stk = C.c_uint32(), C.c_uint32(), C.c_uint32()
    gpu.call("mgv_pool_stats", pool, C.byref(av), C.byref(pr), C.byref(stk), None)
    assert (av.value, stk.value) == (nbuf, nbuf)          # conservation: every buffer back
    gpu.call("mgv_pool_unregister_producer", pool, None)
    gpu.call("mgv_pool_destroy", pool)


def test_select_to_pool_buffer(gpu, orc):
    pool = C.c_void_p()
    gpu.call("mgv_pool_create", 2, 4096, C.byref(pool))
    b = C.c_int32()
    gpu.call("mgv_pool_register_producer", pool, None)
    gpu.call("mgv_pool_acquire", pool, C.byref(b), None)
    n = 6000
    x = orc.gen_range(n, orc.seed(0, 0), 0, 10000) + 1
    d = dev(x)
    gpu.call("mgv_select_to_buffer", d.vp, d.vp, n, 3000, pool, b.value, None)
    mgv.sync()
    data, cnt = C.c_void_p(), C.c_void_p()
    gpu.call("mgv_pool_buffer", pool, b.value, C.byref(data), C.byref(cnt))
    c = np.empty(1, np.uint32)
    gpu.call("mgv_memcpy", c.ctypes.data_as(C.c_void_p), cnt, 4)
    want = orc.select(x, 3000)
    assert c[0] == len(want) and len(want) <= 4096
    out = np.empty(c[0], np.int32)
    gpu.call("mgv_memcpy", out.ctypes.data_as(C.c_void_p), data, 4 * int(c[0]))
    assert np.array_equal(np.sort(out), np.sort(want))
    gpu.call("mgv_pool_destroy", pool)


def test_result_channel(gpu):
    ch = C.c_void_p()
    gpu.call("mgv_channel_create", 4, C.byref(ch))
    res = mgv.ResultBuffer()
    r = mgv.Result()
    seq = C.c_uint32()
    assert mgv.load().mgv_channel_poll(ch, 2, C.byref(r), C.byref(seq)) == 1
    x = np.arange(1000, dtype=np.int32)
    d = dev(x)
    gpu.call("mgv_select_sum", d.vp, d.vp, 1000, 499, res.vp, mgv.launch(), None)
    gpu.call("mgv_channel_emit", ch, 2, res.vp, 77, None)
    t0 = time.time()
    while mgv.load().mgv_channel_poll(ch, 2, C.byref(r), C.byref(seq)) == 1:
        assert time.time() - t0 < 10
    assert (r.count, r.sum, seq.value) == (500, sum(range(500)), 77)
    assert mgv.load().mgv_channel_poll(ch, 2, C.byref(r), C.byref(seq)) == 1
    gpu.call("mgv_channel_destroy", ch)


def test_routed_batch_equals_fused_and_replays(gpu, orc):
    """A C4-shaped batch whose large probes are routed through LDS table
    slices on side streams (forked and joined inside one hipGraph) gives
    every query the fused batch's result, on every replay, and the oracle's
    for the queries the oracle checks."""
    import workloads
    shapes = workloads.c4_shapes(40, seed=0x4C, lo=12.0, hi=20.0)
    st = mgv.Stream()
    batch = workloads.JoinBatch(shapes, 0.25, st.vp)
    st.sync()
    L = mgv.launch(0, 256, 0)
    g = batch.capture_fused(4096, L, st.vp)
    gpu.call("mgv_graph_launch", g, st.vp)
    st.sync()
    fused = batch.results()
    gpu.call("mgv_graph_destroy", g)
    assert int(fused["build_errors"].sum()) == 0
    thr = 1 << 16
    big = [i for i, x in enumerate(shapes) if x.P >= thr and batch.lds_fanout(x)]
    assert 4 <= len(big) < len(shapes)
    g = batch.capture_routed(4096, L, st.vp, threshold=thr, nstreams=3)
    for rep in range(3):
        gpu.call("mgv_graph_launch", g, st.vp)
        st.sync()
        r = batch.results()
        assert np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"]), rep
        batch.res.clear(st.vp)   # eager work between replays
        st.sync()
    gpu.call("mgv_graph_destroy", g)
    # the same routing issued eagerly
    batch.issue_routed(4096, L, st.vp, threshold=thr, nstreams=3)
    st.sync()
    r = batch.results()
    assert np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"])
    for i in big[:3] + [j for j in range(len(shapes)) if j not in big][:3]:
        x, h = shapes[i], batch.host_columns(i)
        tb = orc.table(h["sk"], h["sw"])
        want = orc.probe(tb, h["k"], h["v"], pred=h["c"], bound=x.bound)
        orc.free(tb)
        assert (int(r["count"][i]), int(r["sum"][i])) == (want["count"], want["sum"]), i
    batch.close()
    st.close()
