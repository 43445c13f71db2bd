"""GPU tests of the C++ operator layer (libmgv_ops.so): the reference's plans
(generator -> exchange -> select / hashjoin / aggregation -> gpu_to_cpu ->
exchange -> materializer, src/main9.cu:150-367) built from the drop-in
operators, with device-resident and host-resident columns, compared with the
oracle; and the plain per-vector path of a query batch compared with the
fused batch kernel."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1


@pytest.fixture(scope="module")
def ops(gpu):
    mgv.load_ops()
    mgv.ops_call("mgv_ops_init", 8, 2, 4)
    yield mgv
    mgv.ops_call("mgv_ops_shutdown")


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def host_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def cols(where, *arrays):
    """Column pointers in device memory (where=0) or host memory (where=-1);
    the returned keep-alive list holds the buffers."""
    if where == 0:
        bufs = [dev(a) for a in arrays]
        return [b.vp for b in bufs], bufs
    arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in arrays]
    return [host_ptr(a) for a in arrs], arrs


@pytest.mark.parametrize("where", [0, -1])
@pytest.mark.parametrize("n,vector_rows,block", [(1 << 20, 4096, 256), (300001, 65536, 1024), (1000, 256, 64)])
def test_sum_select_plan(ops, orc, where, n, vector_rows, block):
    a = orc.gen_bits(n, orc.seed(1, 0), 0, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 0, 31)
    bound = (1 << 30) - 1
    (pb, pa), keep = cols(where, b, a)
    dst = np.zeros(4, np.int32)
    s = C.c_int64()
    mgv.ops_call("mgv_ops_sum_select", pb, pa, n, bound, vector_rows, where, 64, block, host_ptr(dst), C.byref(s))
    cnt, want, want32 = orc.select_sum(b, a, bound)
    assert s.value == want
    assert dst[0] == want32          # the int32 value the reference's aggregation delivers
    del keep


def test_sum_select_single_column_and_empty(ops, orc):
    n = 70000
    x = orc.gen_range(n, orc.seed(1, 2), 0, 10000) + 1
    (px,), keep = cols(0, x)
    s = C.c_int64()
    mgv.ops_call("mgv_ops_sum_select", px, px, n, 5000, 4096, 0, 32, 256, None, C.byref(s))
    assert s.value == orc.select_sum(x, x, 5000)[1]
    mgv.ops_call("mgv_ops_sum_select", px, px, 0, 5000, 4096, 0, 32, 256, None, C.byref(s))
    assert s.value == 0
    del keep


@pytest.mark.parametrize("where", [0, -1])
@pytest.mark.parametrize("apply_select", [1, 0])
def test_select_plan_multiset(ops, orc, where, apply_select):
    n = 200003
    a = orc.gen_bits(n, orc.seed(1, 0), 1, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 1, 31)
    bound = (1 << 29) - 1
    (pb, pa), keep = cols(where, b, a)
    dst = np.full(n, -7, np.int32)
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_select", pb, pa, n, bound, apply_select, 8192, where, 64, 256, host_ptr(dst), n,
                 C.byref(cnt))
    want = a[b <= bound] if apply_select else a
    assert cnt.value == len(want)
    assert np.array_equal(np.sort(dst[:cnt.value]), np.sort(want))
    assert np.all(dst[cnt.value:] == -7)
    del keep


# host-resident vectors are copied into the manager's device buffers, which
# this module's mgv_ops_init makes too small for 4 Mi-row vectors
@pytest.mark.parametrize("n,vector_rows,where", [(200003, 8192, 0), (200003, 8192, -1), (9000000, 65536, 0),
                                                 (9000000, 65536, -1), (5000000, 4 << 20, 0), (0, 4096, 0),
                                                 (0, 4096, -1)])
def test_select_split_plan_multiset(ops, orc, n, vector_rows, where):
    """create_select_pipeline_exp: select -> split -> union_all (both split
    parents) -> gpu_to_cpu gives the select's rows; 9e6 rows of 65536-row
    vectors overflow the split's 4 Mi-row staging twice, so its parents get
    flushed vectors mid-stream as well as at close."""
    a = orc.gen_bits(n, orc.seed(1, 4), 2, 31) - (1 << 29)     # negative values too
    b = orc.gen_bits(n, orc.seed(1, 5), 2, 31)
    bound = (1 << 30) - 1
    (pb, pa), keep = cols(where, b, a)
    dst = np.full(n + 1, -7, np.int32)
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_select_split", pb, pa, n, bound, vector_rows, where, 64, 256, host_ptr(dst), n + 1,
                 C.byref(cnt))
    want = a[b <= bound]
    assert cnt.value == len(want)
    assert np.array_equal(np.sort(dst[:cnt.value]), np.sort(want))
    assert np.all(dst[cnt.value:] == -7)
    del keep


@pytest.mark.parametrize("n,vector_rows", [(300001, 4096), (6000000, 1 << 20)])
def test_split_sums_plan(ops, orc, n, vector_rows):
    """select -> split -> {aggregation -> gpu_to_cpu} x 2: each parent sums
    the selected rows of its parity, as the oracle's split routes them."""
    a = orc.gen_bits(n, orc.seed(1, 6), 2, 31) - (1 << 30)
    b = orc.gen_bits(n, orc.seed(1, 7), 2, 31)
    bound = (1 << 30) - 1
    (pb, pa), keep = cols(0, b, a)
    sums = np.zeros(2, np.int64)
    counts = np.zeros(2, np.uint64)
    mgv.ops_call("mgv_ops_split_sums", pb, pa, n, bound, vector_rows, 0, 64, 1024, host_ptr(sums),
                 host_ptr(counts))
    parts = orc.split([a], 2, pred=b, bound=bound)
    assert counts.tolist() == [len(parts[0][0]), len(parts[1][0])]
    assert sums.tolist() == [int(parts[p][0].astype(np.int64).sum()) for p in range(2)]
    del keep


@pytest.mark.parametrize("nstreams,n,vector_rows", [(1, 1 << 20, 8192), (2, (1 << 21) + 77, 8192),
                                                     (4, 3000000, 65536), (3, 0, 4096)])
def test_fan_in_sum_plan(ops, orc, nstreams, n, vector_rows):
    """Several exchange consumers (own streams), one select each, into one
    union_all -> aggregation -> gpu_to_cpu: the union_all's children share
    its result, so the sum is the whole input's."""
    a = orc.gen_bits(n, orc.seed(1, 8), 3, 20)
    b = orc.gen_bits(n, orc.seed(1, 9), 3, 31)
    bound = (1 << 30) - 1
    (pb, pa), keep = cols(0, b, a)
    s = C.c_int64()
    mgv.ops_call("mgv_ops_sum_select_streams", pb, pa, n, bound, vector_rows, 0, 64, 256, nstreams, C.byref(s))
    assert s.value == (orc.select_sum(b, a, bound)[1] if n else 0)
    del keep


@pytest.mark.parametrize("nstreams,where", [(2, 0), (3, -1)])
def test_fan_in_select_plan(ops, orc, nstreams, where):
    """Selected rows of several consumers gathered through one union_all ->
    gpu_to_cpu: the multiset of the select's rows."""
    n = 1500001
    a = orc.gen_bits(n, orc.seed(1, 10), 3, 31)
    b = orc.gen_bits(n, orc.seed(1, 11), 3, 31)
    bound = (1 << 29) - 1
    (pb, pa), keep = cols(where, b, a)
    dst = np.full(n, -7, np.int32)
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_select_streams", pb, pa, n, bound, 16384, where, 64, 256, nstreams, host_ptr(dst), n,
                 C.byref(cnt))
    want = a[b <= bound]
    assert cnt.value == len(want)
    assert np.array_equal(np.sort(dst[:cnt.value]), np.sort(want))
    assert np.all(dst[cnt.value:] == -7)
    del keep


def test_fan_in_rejects_bad_stream_count(ops):
    x = dev(np.arange(1024, dtype=np.int32))
    s = C.c_int64()
    for k in (0, 17):
        with pytest.raises(mgv.MgvError):
            mgv.ops_call("mgv_ops_sum_select_streams", x.vp, x.vp, 1024, 5, 256, 0, 8, 256, k, C.byref(s))


@pytest.mark.parametrize("npipelines,aggregate,where", [(1, 1, 0), (2, 1, 0), (3, 1, -1), (2, 0, 0), (3, 0, -1)])
def test_select_pipelines_with_union_all_cpu(ops, orc, npipelines, aggregate, where):
    """One chain per launch conf joined by mem_move -> union_all_cpu, as the
    reference's plans do for several confs: the materializer sees every
    pipeline's rows (or partial sums) exactly once."""
    n = 1000003
    a = orc.gen_bits(n, orc.seed(1, 12), 4, 31) - (1 << 30)
    b = orc.gen_bits(n, orc.seed(1, 13), 4, 31)
    bound = (1 << 29) - 1
    (pb, pa), keep = cols(where, b, a)
    cap = n if not aggregate else npipelines
    dst = np.full(cap, -7, np.int32)
    v = C.c_int64()
    mgv.ops_call("mgv_ops_select_pipelines", pb, pa, n, bound, 16384, where, 64, 256, npipelines, aggregate,
                 host_ptr(dst), cap, C.byref(v))
    want = a[b <= bound]
    if aggregate:
        assert v.value == int(want.astype(np.int64).sum())
        # one low-32-bit partial per pipeline; their int32 sum wraps like the exact sum's low bits
        assert (int(dst.astype(np.int64).sum()) - v.value) % (1 << 32) == 0
    else:
        assert v.value == len(want)
        assert np.array_equal(np.sort(dst[:v.value]), np.sort(want))
        assert np.all(dst[v.value:] == -7)
    del keep


def pool_stats():
    v = [C.c_uint32() for _ in range(4)]
    mgv.ops_call("mgv_ops_pool_stats", *[C.byref(x) for x in v])
    return [x.value for x in v]


def test_select_plan_streams_more_rows_than_the_device_pool(ops, orc):
    """select -> gpu_to_cpu with more survivors than the manager's device pool
    holds at once (8 buffers of 4 Mi rows): the kernels rotate the pool's
    buffers (acquire_buffer_blocked_to, src/buffer_manager.cuh:147-176), the
    catcher hands each full one back, and at the end every device buffer is
    back in the pool or the manager's cache."""
    n = (1 << 26) + 4097
    a = orc.gen_bits(n, orc.seed(1, 0), 50, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 50, 31)
    bound = int(0.52 * 2**31)
    want = np.sort(a[b <= bound])
    free0, cached0, total, _ = pool_stats()
    assert len(want) > total * (4 << 20)
    (pb, pa), keep = cols(0, b, a)
    dst = np.full(n, -7, np.int32)
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_select", pb, pa, n, bound, 1, 1 << 20, 0, 128, 1024, host_ptr(dst), n, C.byref(cnt))
    assert cnt.value == len(want)
    assert np.array_equal(np.sort(dst[:cnt.value]), want)
    assert np.all(dst[cnt.value:] == -7)
    free, cached, total2, _ = pool_stats()
    assert total2 == total and free + cached == total, (free, cached, total)
    del keep


def test_select_plan_rejects_small_dst(ops):
    x = dev(np.arange(1024, dtype=np.int32))
    dst = np.zeros(10, np.int32)
    cnt = C.c_uint64()
    with pytest.raises(mgv.MgvError) as e:
        mgv.ops_call("mgv_ops_select", x.vp, x.vp, 1024, 5, 1, 256, 0, 8, 256, host_ptr(dst), 10, C.byref(cnt))
    assert e.value.code == mgv.E_CAPACITY


@pytest.mark.parametrize("where", [0, -1])
@pytest.mark.parametrize("with_select", [False, True])
def test_count_join_plan(ops, orc, where, with_select):
    nb, npr = 1 << 16, 1 << 20
    R = orc.gen_perm(nb, nb, orc.seed(2, 0), 5)
    L = orc.gen_range(npr, orc.seed(2, 2), 5, 2 * nb)
    pr = orc.gen_bits(nb, orc.seed(3, 0), 5, 31)
    pl = orc.gen_bits(npr, orc.seed(3, 1), 5, 31)
    bound = (1 << 30) - 1
    (dR, dL, dpr, dpl), keep = cols(where, R, L, pr, pl)
    cnt = C.c_uint64()
    dst = np.zeros(2, np.int32)
    mgv.ops_call("mgv_ops_count_join", dR, dpr if with_select else None, nb, dL, dpl if with_select else None, npr,
                 bound, 16384, where, 128, 256, host_ptr(dst), C.byref(cnt))
    if with_select:
        m = pr <= bound
        t = orc.table(np.ascontiguousarray(R[m]))
        want = orc.probe(t, L, pred=pl, bound=bound)
    else:
        t = orc.table(R)
        want = orc.probe(t, L)
    orc.free(t)
    assert cnt.value == want["count"]
    assert dst[0] == want["count32"]
    del keep


@pytest.mark.parametrize("with_select", [False, True])
def test_count_join_plan_large_tables_go_through_lds_slices(ops, orc, with_select):
    """create_count_join_1G_pipeline at 2^24 build keys (a 256 MiB table) and
    2^26 probe rows in 4 Mi-row vectors: the drop-in operators build and
    probe those vectors through LDS table slices; the count is the closed
    form #{L < 2^24} (keys a permutation of 0..2^24-1), with the selects
    applied to both sides in the second case."""
    nb, npr = 1 << 24, 1 << 26
    R = orc.gen_perm(nb, nb, orc.seed(2, 0), 6)
    L = orc.gen_bits(npr, orc.seed(2, 2), 6, 25)
    bound = (1 << 30) - 1
    (dR, dL), keep = cols(0, R, L)
    if with_select:
        pr = orc.gen_bits(nb, orc.seed(3, 0), 6, 31)
        pl = orc.gen_bits(npr, orc.seed(3, 1), 6, 31)
        (dpr, dpl), keep2 = cols(0, pr, pl)
        inv = np.empty(nb, np.int64)
        inv[R] = np.arange(nb)
        hit = L < nb
        ok = np.zeros(npr, bool)
        ok[hit] = pr[inv[L[hit]]] <= bound
        want = int((ok & (pl <= bound)).sum())
    else:
        dpr = dpl = None
        want = int((L < nb).sum())
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_count_join", dR, dpr, nb, dL, dpl, npr, bound, 4 << 20, 0, 256, 1024, None, C.byref(cnt))
    assert cnt.value == want
    del keep


def test_count_join_plan_with_duplicate_keys(ops, orc):
    """Self-join closed form: sum over values v <= 5000 of hist[v]^2 (main7.cu:62-68)."""
    n = 100000
    x = orc.gen_range(n, orc.seed(2, 5), 0, 10000) + 1
    (px,), keep = cols(0, x)
    pred = dev(x)
    cnt = C.c_uint64()
    mgv.ops_call("mgv_ops_count_join", px, pred.vp, n, px, None, n, 5000, 4096, 0, 64, 256, None, C.byref(cnt))
    assert cnt.value == orc.selfjoin_count(x, 5000, 10000)
    del keep


def test_plain_path_equals_fused_batch(ops, orc):
    """Each query of a batch gives the same result through one launch per
    vector (the plain path) and through the fused persistent kernel."""
    from test_gpu_batch_pool import make_queries, run_batch
    qs = make_queries(orc, 10, seed_cfg=21)
    fused, _, _ = run_batch(mgv, qs, 1024, 1, 256)
    nq = len(qs)
    tables = [mgv.HashTable(q["B"], 0.5) for q in qs]
    res = mgv.ResultBuffer(2 * nq)
    qarr = (mgv.Query * (2 * nq))()
    for i, (q, t) in enumerate(zip(qs, tables)):
        d = q["d"]
        qarr[i] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, d["sk"].ptr, d["sw"].ptr, q["B"], t.buf.ptr, t.nbuckets, 0,
                            None, None, res.ptr + 64 * i, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
        qarr[nq + i] = mgv.Query(mgv.MGV_Q_PROBE_SUM, q["bound"], d["c"].ptr, d["rk"].ptr, d["rv"].ptr, q["P"],
                                 t.buf.ptr, t.nbuckets, 0, None, None, res.ptr + 64 * (nq + i), 0, 0, 0)
    mgv.ops_call("mgv_ops_run_plain", qarr, 2 * nq, 1024, mgv.launch(64, 256), None)
    mgv.sync()
    for i, q in enumerate(qs):
        r = res.read(nq + i)
        assert res.read(i).errors == 0
        assert (r.count, r.sum) == (fused[i].count, fused[i].sum) == (q["want"]["count"], q["want"]["sum"])


def test_plan_failure_stops_the_generator(ops):
    """A consumer whose launch fails (misaligned column, rejected on the host
    before any kernel runs) ends the plan with that error instead of leaving
    the generator blocked on a free buffer; the manager gets its buffers back."""
    x = dev(np.arange(1 << 17, dtype=np.int32))
    bad = C.c_void_p(x.ptr + 4)
    s = C.c_int64()
    for _ in range(2):
        with pytest.raises(mgv.MgvError) as e:
            mgv.ops_call("mgv_ops_sum_select", bad, bad, (1 << 17) - 64, 5, 4096, 0, 8, 256, None, C.byref(s))
        assert e.value.code == mgv.E_ALIGN
    mgv.ops_call("mgv_ops_sum_select", x.vp, x.vp, 1 << 17, 5, 4096, 0, 8, 256, None, C.byref(s))
    assert s.value == sum(range(6))
