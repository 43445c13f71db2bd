"""GPU tests of the continuous-batching serving loop (mgv_serve_run): each
request's result is the same alone (one request per batch, one batch at a
time) and in the mix (Poisson arrivals, batches of many requests, several
batches in flight), and equals the oracle; every result reaches the host
through the device-written completion ring; admission respects the HBM
budget."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(gpu):
    mgv.load_ops()
    t = workloads.ServeTables(build_logs=range(8, 15), probe_cap=1 << 17, group_rows=(1 << 18, 3 << 16),
                              ngroups=1 << 12, zipf_s=0.8)
    mgv.sync()
    yield t
    t.close()


def oracle_results(orc, tables, desc):
    out = []
    for d in desc:
        if d["kind"] == "join":
            p = tables.pairs[d["pair"]]
            B, P = p["B"], p["P"]
            sk, sw = p["sk"].download(np.int32, B), p["sw"].download(np.int32, B)
            c, k, v = (p[x].download(np.int32, P) for x in ("rc", "rk", "rv"))
            t = orc.table(sk, sw)
            w = orc.probe(t, k, v, pred=c, bound=d["bound"])
            orc.free(t)
            out.append((w["count"], w["sum"], 0))
        else:
            g = tables.groups[d["table"]]
            k, v = g["k"].download(np.int32, d["rows"]), g["v"].download(np.int32, d["rows"])
            cnt, s = orc.groupby(k, v, tables.ngroups)
            out.append((int(cnt.sum()), int(s.sum()), workloads.group_digest(cnt, s)))
    return out


def triples(res, n):
    return [(res[i].count, res[i].sum, res[i].digest) for i in range(n)]


def test_serve_mix_equals_alone_equals_oracle(tables, orc):
    n = 48
    reqs, desc = tables.requests(n, rate_qps=20000.0, seed=3)
    mix, st = workloads.serve(reqs, n, 256 << 20, max_inflight=4, max_batch=16, vector_rows=4096)
    alone, st1 = workloads.serve(reqs, n, 256 << 20, max_inflight=1, max_batch=1, vector_rows=4096)
    assert st1.batches == n and st1.max_inflight_seen == 1
    assert triples(mix, n) == triples(alone, n)
    assert triples(mix, n) == oracle_results(orc, tables, desc)
    assert all(mix[i].errors == 0 for i in range(n))
    assert all(mix[i].latency_s > 0 and mix[i].done_s >= mix[i].admit_s for i in range(n))
    assert st.p99_s >= st.p50_s > 0 and st.queries_per_s > 0
    assert st.batches < n                      # requests were batched together
    # every result reached the host through the device-written completion ring
    assert st.ring_results == n and st1.ring_results == n


def test_serve_budget_limits_admission(tables, orc):
    """A budget that holds only the largest request's table: requests run
    with at most that much reserved, and the results do not change."""
    n = 24
    reqs, desc = tables.requests(n, rate_qps=0.0, seed=5, arrivals=False)
    biggest = max(mgv.ht_plan(d["B"], 0.5)[1] if d["kind"] == "join" else tables.ngroups * 16 for d in desc)
    budget = (biggest + 255) // 256 * 256
    res, st = workloads.serve(reqs, n, budget, max_inflight=4, max_batch=8, vector_rows=4096)
    assert st.peak_bytes <= budget
    ref, _ = workloads.serve(reqs, n, 256 << 20, max_inflight=4, max_batch=8, vector_rows=4096)
    assert triples(res, n) == triples(ref, n)


def test_serve_rejects_request_above_budget(tables):
    reqs, desc = tables.requests(4, rate_qps=0.0, seed=7, arrivals=False, join_frac=1.0)
    with pytest.raises(mgv.MgvError) as e:
        workloads.serve(reqs, 4, 1024, max_inflight=1, max_batch=1, vector_rows=4096)
    assert e.value.code == mgv.E_CAPACITY


def test_serve_rejects_unsorted_arrivals(tables):
    reqs, _ = tables.requests(4, rate_qps=1000.0, seed=9)
    reqs[0].arrival_s, reqs[3].arrival_s = reqs[3].arrival_s + 1.0, reqs[0].arrival_s
    with pytest.raises(mgv.MgvError) as e:
        workloads.serve(reqs, 4, 256 << 20)
    assert e.value.code == mgv.E_INVALID


def test_serve_many_groups_use_partitioned_groupby(gpu, orc):
    """Group-bys over more than 4096 groups run as the partitioned group-by
    inside the loop; results equal the oracle alone and in the mix."""
    t = workloads.ServeTables(build_logs=range(8, 10), probe_cap=1 << 14, group_rows=(1 << 19, 5 << 16),
                              ngroups=1 << 14, zipf_s=0.8)
    try:
        n = 16
        reqs, desc = t.requests(n, rate_qps=50000.0, seed=11, join_frac=0.25, min_group_log=12)
        mix, st = workloads.serve(reqs, n, 512 << 20, max_inflight=3, max_batch=8, vector_rows=4096)
        alone, _ = workloads.serve(reqs, n, 512 << 20, max_inflight=1, max_batch=1, vector_rows=4096)
        assert triples(mix, n) == triples(alone, n) == oracle_results(orc, t, desc)
        big = [i for i, d in enumerate(desc) if d["kind"] == "groupby"]
        assert big and all(mix[i].footprint > (1 << 14) * 16 for i in big)   # workspace counted
    finally:
        t.close()


def test_serve_large_builds_through_lds_slices(gpu, orc):
    """Joins whose build side has >= 2^21 rows build their table through LDS
    slices (their footprint then includes the build workspace); results are
    the same alone and in the mix, and equal the oracle."""
    mgv.load_ops()
    t = workloads.ServeTables(build_logs=(20, 21, 22), probe_cap=1 << 20, group_rows=(1 << 16,),
                              ngroups=1 << 10, zipf_s=0.8)
    mgv.sync()
    try:
        n = 12
        reqs, desc = t.requests(n, rate_qps=5000.0, seed=5)
        mix, _ = workloads.serve(reqs, n, 1 << 30, max_inflight=3, max_batch=4, vector_rows=4096)
        alone, _ = workloads.serve(reqs, n, 1 << 30, max_inflight=1, max_batch=1, vector_rows=4096)
        assert triples(mix, n) == triples(alone, n) == oracle_results(orc, t, desc)
        big = [i for i, d in enumerate(desc) if d["kind"] == "join" and t.pairs[d["pair"]]["B"] >= 1 << 21]
        assert big, "the request mix has no large build"
        for i in big:
            _, table_bytes = mgv.ht_plan(t.pairs[desc[i]["pair"]]["B"], 0.5)
            assert mix[i].footprint > table_bytes   # table + LDS-build workspace
        small = [i for i, d in enumerate(desc) if d["kind"] == "join" and t.pairs[d["pair"]]["B"] < 1 << 21]
        for i in small:
            _, table_bytes = mgv.ht_plan(t.pairs[desc[i]["pair"]]["B"], 0.5)
            assert mix[i].footprint <= (table_bytes + 255) // 256 * 256
    finally:
        t.close()
