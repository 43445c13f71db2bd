"""GPU test of the autotuner in quick mode on reduced workloads: every
candidate of every search space must reproduce the first candidate's
results (the tuner raises otherwise), the C4 space covers the five
workgroup sizes and the routed fan-outs (the batched 256-query workload's
space: vector size, workgroup 64..1024, load factor, partition fan-out),
and the C2 space every probe form."""
from __future__ import annotations

import pytest

import autotune
import mgv

pytestmark = pytest.mark.gpu


def test_autotune_quick_spaces_keep_results_identical(gpu):
    st = mgv.Stream()
    c4 = autotune.tune_c4(st, quick=True, nq=24)
    cands = c4["candidates"]
    assert len({c["digest"] for c in cands}) == 1   # fused, routed and LDS-batch candidates alike
    launches = [c for c in cands if "block" in c]
    lds = [c for c in cands if "min_rows" in c]
    assert lds and len(launches) + len(lds) == len(cands)
    assert {c["block"] for c in launches} == {64, 128, 256, 512, 1024}
    routed = [c for c in launches if c["route_fanout"] > 1]
    assert routed and all(c["routed_queries"] > 0 for c in routed)
    assert c4["route_fanout"] in {c["route_fanout"] for c in launches}
    assert c4["ms"] == min(c["ms"] for c in cands)
    assert c4["path"] in ("fused", "routed", "lds_batch")
    assert c4["lds_batch"]["ms"] == min(c["ms"] for c in lds)
    assert 1024 in {c["vector_rows"] for c in lds}   # the direct queries' own vector size

    c2 = autotune.tune_c2(st, quick=True, ns=1 << 20, nr=1 << 22)
    assert len({c["digest"] for c in c2["candidates"]}) == 1
    assert {c["mode"] for c in c2["candidates"]} == {"direct", "part", "lds", "ldsk", "ldsc"}

    c1 = autotune.tune_c1(st, quick=True, n=1 << 22)
    assert len({c["digest"] for c in c1["candidates"]}) == 1
    st.close()
