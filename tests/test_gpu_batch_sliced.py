"""GPU tests of the XCD-sliced batch probe (mgv_batch_plan_sliced /
mgv_batch_run_sliced): every query's COUNT and SUM equal the oracle's and the
fused batch's, with all, some or none of the queries sliced, every launch
shape, skewed probe keys that overfill a region (rows probed on the spot),
duplicate build keys, and on every replay of a captured graph."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

INT_MIN, INT_MAX = -2**31, 2**31 - 1


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def make_queries(orc, nq, seed_cfg=11, lo_log=8, hi_log=17, hot=None, dup=()):
    """select (c <= bound) -> join (k) -> SUM(v + w) queries with their
    oracle results.  hot: query index whose probe keys are 90 % one key;
    dup: query indices whose build keys repeat (about 4 times each)."""
    rng = np.random.default_rng(seed_cfg)
    qs = []
    for q in range(nq):
        P = int(2 ** rng.uniform(lo_log, hi_log)) + int(rng.integers(0, 7))
        B = max(16, P // 16)
        sk = orc.gen_range(B, orc.seed(9, 3), q, max(1, B // 4)) if q in dup else orc.gen_perm(B, B, orc.seed(9, 3), q)
        sw = orc.gen_bits(B, orc.seed(9, 4), q, 16)
        rk = orc.gen_range(P, orc.seed(9, 1), q, 2 * B)
        if q == hot:
            rk[: (9 * P) // 10] = sk[0]
        rv = orc.gen_bits(P, orc.seed(9, 2), q, 16)
        c = orc.gen_bits(P, orc.seed(9, 0), q, 31)
        bound = int(rng.integers(INT_MIN // 2, INT_MAX)) if q != hot else INT_MAX
        t = orc.table(sk, sw)
        want = orc.probe(t, rk, rv, pred=c, bound=bound)
        orc.free(t)
        qs.append(dict(P=P, B=B, bound=bound, want=want,
                       d={k: dev(v) for k, v in dict(sk=sk, sw=sw, rk=rk, rv=rv, c=c).items()}))
    return qs


def upload(arr):
    return mgv.DeviceBuffer(max(16, C.sizeof(arr))).upload(np.frombuffer(bytes(arr), dtype=np.uint8))


def build_tables(gpu, qs, lf=0.5, lds_min_rows=None, fresh=False):
    """The queries' tables, built by one fused launch, or (lds_min_rows set)
    by the LDS batch build (builds of at least lds_min_rows rows whose table
    splits into 1024-bucket slices; the others through the fused launch).
    fresh: the LDS-built tables are first filled with a non-empty pattern
    and built with MGV_LAUNCH_FRESH_TABLES, which must overwrite it."""
    nq = len(qs)
    tables = [mgv.HashTable(q["B"], lf) for q in qs]
    res = mgv.ResultBuffer(2 * nq)   # [0, nq): builds, [nq, 2nq): probes
    build_q = (mgv.Query * nq)()
    probe_q = (mgv.Query * nq)()
    for i, (q, t) in enumerate(zip(qs, tables)):
        d = q["d"]
        build_q[i] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, d["sk"].ptr, d["sw"].ptr, q["B"], t.buf.ptr, t.nbuckets, 0,
                               None, None, res.ptr + 64 * i, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
        probe_q[i] = mgv.Query(mgv.MGV_Q_PROBE_SUM, q["bound"], d["c"].ptr, d["rk"].ptr, d["rv"].ptr, q["P"],
                               t.buf.ptr, t.nbuckets, 0, None, None, res.ptr + 64 * (nq + i), 0, 0, 0)
    if lds_min_rows is None:
        n = C.c_uint64()
        gpu.call("mgv_batch_plan", build_q, nq, 1024, None, 0, C.byref(n))
        items = (mgv.WorkItem * n.value)()
        gpu.call("mgv_batch_plan", build_q, nq, 1024, items, n.value, C.byref(n))
        dq, di = upload(build_q), upload(items)
        counter = mgv.DeviceBuffer(64)
        gpu.call("mgv_batch_run", dq.vp, di.vp, n.value, counter.vp, mgv.launch(0, 256, 1), None)
    else:
        plan = mgv.LdsBatchPlan()
        regions = (C.c_uint64 * (4 * nq))()
        gpu.call("mgv_batch_plan_lds", build_q, nq, 1024, lds_min_rows, None, 0, regions, C.byref(plan))
        items = (mgv.WorkItem * max(1, plan.nitems))()
        gpu.call("mgv_batch_plan_lds", build_q, nq, 1024, lds_min_rows, items, plan.nitems, regions, C.byref(plan))
        assert plan.nparts == 0
        dq, di, dr = upload(build_q), upload(items), upload(regions)
        ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
        L = mgv.launch(0, 256, 1)
        if fresh:
            built = [i for i in range(nq) if regions[4 * i + 2]]
            assert built
            for i in built:
                gpu.call("mgv_fill_u64", tables[i].buf.vp, tables[i].nbuckets * 8, 0x0123456789ABCDEF, None)
            L = mgv.Launch(0, 256, 1, mgv.MGV_LAUNCH_FRESH_TABLES)
        gpu.call("mgv_batch_run_lds", dq.vp, di.vp, dr.vp, C.byref(plan), ws.vp, ws.nbytes, L, None)
    mgv.sync()
    assert all(res.read(i).errors == 0 for i in range(nq))
    return tables, res, probe_q


def plan_sliced(gpu, probe_q, nq, vector_rows, min_rows, nregions=0):
    plan = mgv.SlicedPlan()
    regions = (C.c_uint64 * (4 * nq))()
    gpu.call("mgv_batch_plan_sliced", probe_q, nq, vector_rows, min_rows, nregions, None, 0, regions, C.byref(plan))
    items = (mgv.WorkItem * max(1, plan.nitems))()
    gpu.call("mgv_batch_plan_sliced", probe_q, nq, vector_rows, min_rows, nregions, items, plan.nitems, regions,
             C.byref(plan))
    return plan, items, regions


COOP = 1   # MGV_LAUNCH_COOP_PROBE: 4 lanes per row in the region probe


def run_sliced(gpu, qs, vector_rows, min_rows, block=256, compact=1, flags=0, nregions=0):
    nq = len(qs)
    tables, res, probe_q = build_tables(gpu, qs)
    plan, items, regions = plan_sliced(gpu, probe_q, nq, vector_rows, min_rows, nregions)
    dq, di, dr = upload(probe_q), upload(items), upload(regions)
    ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
    gpu.call("mgv_batch_run_sliced", dq.vp, di.vp, dr.vp, C.byref(plan), ws.vp, ws.nbytes,
             mgv.Launch(0, block, compact, flags), None)
    mgv.sync()
    sliced = sum(1 for i in range(nq) if regions[4 * i + 2])
    return [res.read(nq + i) for i in range(nq)], sliced, plan


@pytest.mark.parametrize("vector_rows,min_rows,block,compact,flags,nregions",
                         [(1024, 0, 256, 1, 0, 32), (4096, 1 << 12, 64, 0, COOP, 8), (256, 1 << 30, 1024, 1, 0, 0),
                          (2048, 0, 512, 0, COOP, 16), (8192, 1 << 14, 128, 1, 0, 8), (4096, 0, 1024, 1, COOP, 32)])
def test_sliced_batch_matches_oracle(gpu, orc, vector_rows, min_rows, block, compact, flags, nregions):
    qs = make_queries(orc, 24)
    got, sliced, plan = run_sliced(gpu, qs, vector_rows, min_rows, block, compact, flags, nregions)
    assert plan.nregions == (nregions or 32)
    want_sliced = sum(1 for q in qs if q["P"] >= min_rows)
    assert sliced == want_sliced
    assert plan.nscatter == sum((q["P"] + 4095) // 4096 for q in qs if q["P"] >= min_rows)
    for q, r in zip(qs, got):
        assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), (q["P"], q["bound"])


def test_sliced_batch_skewed_keys_and_duplicate_build_keys(gpu, orc):
    """Query 3's probe rows are 90 % one key, so one region overflows and
    its excess rows are probed by the scatter itself; queries 5 and 9 have
    duplicate build keys (several hitting slots per row)."""
    qs = make_queries(orc, 12, seed_cfg=5, lo_log=14, hi_log=18, hot=3, dup=(5, 9))
    for flags, nregions in ((0, 8), (COOP, 32)):
        got, sliced, _ = run_sliced(gpu, qs, 2048, 0, flags=flags, nregions=nregions)
        assert sliced == len(qs)
        for q, r in zip(qs, got):
            assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), (q["P"], q["bound"], flags)


def test_sliced_batch_equals_fused_on_graph_replays(gpu, orc):
    """A C4-shaped batch (40 queries, probes of 2^12-2^20 rows, load factor
    0.25): the sliced probe launch captured with the builds into one graph
    gives every query the fused batch's result on every replay (eager work
    between replays), and the oracle's for the checked queries."""
    import workloads
    shapes = workloads.c4_shapes(40, seed=0x5C, lo=12.0, hi=20.0)
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
    g = batch.capture_sliced(4096, L, st.vp, min_rows=thr)
    assert 4 <= batch._sliced(4096, thr)["queries"] < len(shapes)
    for rep in range(3):
        gpu.call("mgv_graph_launch", g, st.vp)
        st.sync()
        r = batch.results()
        assert np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"]), rep
        batch.res.clear(st.vp)   # eager work between replays
        st.sync()
    gpu.call("mgv_graph_destroy", g)
    batch.issue_sliced(4096, L, st.vp, min_rows=thr)
    st.sync()
    r = batch.results()
    assert np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"])
    big = [i for i, x in enumerate(shapes) if x.P >= thr]
    for i in big[:3] + [j for j in range(len(shapes)) if j not in big][:2]:
        x, h = shapes[i], batch.host_columns(i)
        tb = orc.table(h["sk"], h["sw"])
        want = orc.probe(tb, h["k"], h["v"], pred=h["c"], bound=x.bound)
        orc.free(tb)
        assert (int(r["count"][i]), int(r["sum"][i])) == (want["count"], want["sum"]), i
    batch.close()
    st.close()


def test_sliced_batch_rejects_bad_arguments(gpu, orc):
    qs = make_queries(orc, 4)
    nq = len(qs)
    tables, res, probe_q = build_tables(gpu, qs)
    plan, items, regions = plan_sliced(gpu, probe_q, nq, 1024, 0)
    dq, di, dr = upload(probe_q), upload(items), upload(regions)
    ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
    L = mgv.launch(0, 256, 1)
    with pytest.raises(mgv.MgvError):   # workspace one byte short
        gpu.call("mgv_batch_run_sliced", dq.vp, di.vp, dr.vp, C.byref(plan), ws.vp, plan.workspace_bytes - 1, L,
                 None)
    bad = mgv.SlicedPlan.from_buffer_copy(plan)
    bad.section[3], bad.section[4] = bad.section[4], bad.section[3] + 1
    with pytest.raises(mgv.MgvError):   # sections out of order
        gpu.call("mgv_batch_run_sliced", dq.vp, di.vp, dr.vp, C.byref(bad), ws.vp, ws.nbytes, L, None)
    with pytest.raises(mgv.MgvError):   # too few items for the plan
        small = (mgv.WorkItem * 1)()
        gpu.call("mgv_batch_plan_sliced", probe_q, nq, 1024, 0, 0, small, 1, regions, C.byref(plan))
    with pytest.raises(mgv.MgvError):   # vector size not a multiple of 256
        gpu.call("mgv_batch_plan_sliced", probe_q, nq, 1000, 0, 0, None, 0, regions, C.byref(plan))
    with pytest.raises(mgv.MgvError):   # region count not 8, 16 or 32
        gpu.call("mgv_batch_plan_sliced", probe_q, nq, 1024, 0, 12, None, 0, regions, C.byref(plan))
    mgv.sync()
    # nothing ran: every probe result is still zero
    assert all((res.read(nq + i).count, res.read(nq + i).sum) == (0, 0) for i in range(nq))


# ---------------------------------------------------------------- LDS batch probe

def run_ldsb(gpu, qs, vector_rows, min_rows, block=256, compact=1, lf=0.5, lds_build=True):
    nq = len(qs)
    tables, res, probe_q = build_tables(gpu, qs, lf, min_rows if lds_build else None)
    plan = mgv.LdsBatchPlan()
    regions = (C.c_uint64 * (4 * nq))()
    gpu.call("mgv_batch_plan_lds", probe_q, nq, vector_rows, min_rows, None, 0, regions, C.byref(plan))
    items = (mgv.WorkItem * max(1, plan.nitems))()
    gpu.call("mgv_batch_plan_lds", probe_q, nq, vector_rows, min_rows, items, plan.nitems, regions, C.byref(plan))
    dq, di, dr = upload(probe_q), upload(items), upload(regions)
    ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
    gpu.call("mgv_batch_run_lds", dq.vp, di.vp, dr.vp, C.byref(plan), ws.vp, ws.nbytes,
             mgv.launch(0, block, compact), None)
    mgv.sync()
    lds = [i for i in range(nq) if regions[4 * i + 2]]
    for i in lds:   # the slices cover the table
        assert regions[4 * i + 2] * regions[4 * i + 3] == tables[i].nbuckets and regions[4 * i + 3] <= 2048
    return [res.read(nq + i) for i in range(nq)], lds, plan


@pytest.mark.parametrize("vector_rows,min_rows,block,compact,lf", [(1024, 0, 256, 1, 0.5), (4096, 1 << 12, 64, 0, 0.25),
                                                                   (2048, 1 << 30, 512, 1, 0.5),
                                                                   (8192, 1 << 14, 1024, 0, 0.75)])
def test_lds_batch_matches_oracle(gpu, orc, vector_rows, min_rows, block, compact, lf):
    qs = make_queries(orc, 24, seed_cfg=13, hi_log=19)
    # the oracle's tables do not depend on the load factor; the GPU tables do
    got, lds, plan = run_ldsb(gpu, qs, vector_rows, min_rows, block, compact, lf)
    assert plan.nparts == 0 or len(lds) > 0
    if min_rows <= 1 << 12:
        assert len(lds) >= 8
    for q, r in zip(qs, got):
        assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), (q["P"], q["bound"])


def test_lds_batch_skewed_keys_and_duplicate_build_keys(gpu, orc):
    """Probes of 2^18.1-2^19.5 rows: their tables (more than 4096 buckets, a
    multiple of 256) all split into slices, so query 3's 90 % of rows on one
    key overfill their region and are probed by the scatter itself; queries
    5 and 9 have duplicate build keys."""
    qs = make_queries(orc, 12, seed_cfg=5, lo_log=18.1, hi_log=19.5, hot=3, dup=(5, 9))
    got, lds, _ = run_ldsb(gpu, qs, 2048, 0)
    assert len(lds) == len(qs)
    for q, r in zip(qs, got):
        assert (r.count, r.sum) == (q["want"]["count"], q["want"]["sum"]), (q["P"], q["bound"])


@pytest.mark.parametrize("fresh", [False, True])
@pytest.mark.parametrize("lf", [0.5, 0.75])
def test_lds_batch_build_equals_fused_build(gpu, orc, lf, fresh):
    """Tables built by the LDS batch build hold exactly the slot multiset of
    the fused build's tables and fill every bucket's slots in order (unique
    and duplicate build keys); with MGV_LAUNCH_FRESH_TABLES also over tables
    that held other slots before the call."""
    qs = make_queries(orc, 10, seed_cfg=17, lo_log=18.1, hi_log=20, dup=(2, 7))
    a, _, _ = build_tables(gpu, qs, lf)
    b, _, _ = build_tables(gpu, qs, lf, lds_min_rows=0, fresh=fresh)
    empty = np.uint64(0xFFFFFFFFFFFFFFFF)
    for ta, tb in zip(a, b):
        sa, sb = ta.slots(), tb.slots()
        assert np.array_equal(np.sort(sa[sa != empty]), np.sort(sb[sb != empty]))
        occ = (sb != empty).reshape(tb.nbuckets, 8)
        assert np.array_equal(occ, np.arange(8)[None, :] < occ.sum(axis=1)[:, None])


def test_lds_batch_equals_fused_on_graph_replays(gpu, orc):
    """The C4-shaped batch of the sliced test through the LDS batch probe,
    captured with the builds into one graph: every query equals the fused
    batch on every replay."""
    import workloads
    shapes = workloads.c4_shapes(40, seed=0x5C, lo=12.0, hi=20.0)
    st = mgv.Stream()
    batch = workloads.JoinBatch(shapes, 0.25, st.vp)
    st.sync()
    L = mgv.launch(0, 256, 0)
    g = batch.capture_fused(4096, L, st.vp)
    gpu.call("mgv_graph_launch", g, st.vp)
    st.sync()
    fused = batch.results()
    gpu.call("mgv_graph_destroy", g)
    g = batch.capture_ldsb(4096, L, st.vp, min_rows=1 << 14)
    assert batch._ldsb(4096, 1 << 14)["queries"] >= 4
    # the LDS-built tables are one span, left uncleared: every replay writes them whole
    assert batch._ldsb(4096, 1 << 14, True)["span"] is not None
    for rep in range(3):
        gpu.call("mgv_graph_launch", g, st.vp)
        st.sync()
        r = batch.results()
        assert np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"]), rep
        batch.res.clear(st.vp)
        st.sync()
    gpu.call("mgv_graph_destroy", g)
    batch.close()
    st.close()
