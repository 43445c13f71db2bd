"""Autotuner of the launch parameters (run on the GPU; results persisted to
tuned.json beside this file, which bench.py reads).

Workloads and search spaces:
  c4  256-query select -> join -> SUM batch in one graph: vector size {1K, 4K,
      16K, 64K} x workgroup {64, 128, 256, 512, 1024} x select compaction
      {0, 1} x hash-table load factor {0.25, 0.5, 0.75}; then, at the best of
      those, the partition fan-out of the queries' probes: 1 (every probe in
      the fused kernel) or the large probes (>= 2^20 rows) routed through
      the LDS probe with level-1 fan-out {16, 64, 256} on side streams of
      the same graph; and the LDS batch (builds and probes of at least
      {2^12, 2^14, 2^16} rows through LDS slices, the other queries fused at
      vector size {1K, the best}) at each load factor
  c2  2^28-row probe of a 2^24-key table (load factor 0.5): the direct probe
      (workgroup {64, 256, 1024} x {per-lane, 4-lane cooperative}), the
      partitioned probe (fan-out {32..256} x scatter shape) and the LDS probe
      (level-1 fan-out {16..128}; tagged slices, key slices, counted
      partitionings)
  c1  filtered-sum scan of 2^28 rows: workgroup x compaction

Every candidate must give results identical to the first candidate's
(bit-exact COUNT / SUM per query, recorded as a digest with each candidate);
a candidate that disagrees is an error, not a slower entry.

  python multigpu-vector_amd/autotune.py --workloads c4,c2,c1
  python multigpu-vector_amd/autotune.py --quick --out /tmp/t.json   # reduced spaces
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import mgv  # noqa: E402
import workloads  # noqa: E402

OUT = os.path.join(HERE, "tuned.json")
MGV_LAUNCH_COUNTED_PART = 1 << 12
MGV_LAUNCH_KEY_SLICES = 1 << 13


class CandidateMismatch(RuntimeError):
    """A candidate's results differ from the reference candidate's."""


def log(*a):
    print("[autotune]", *a, file=sys.stderr, flush=True)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def check_same(ref, got, what: str):
    """The result-equality rule: returns the reference (the first digest
    seen) and raises CandidateMismatch when a later candidate's differs."""
    if ref is not None and got != ref:
        raise CandidateMismatch(f"{what} changed the results ({got} != {ref})")
    return got if ref is None else ref


def time_ms(fn, st, reps=3):
    best = None
    for _ in range(reps):
        a, b = mgv.Event(), mgv.Event()
        a.record(st.vp)
        fn()
        b.record(st.vp)
        st.sync()
        ms = a.elapsed_ms(b)
        a.close()
        b.close()
        best = ms if best is None else min(best, ms)
    return best


def tune_c4(st, quick=False, nq=256):
    shapes = workloads.c4_shapes(nq)
    batch = workloads.JoinBatch(shapes, 0.5, st.vp)
    st.sync()
    Vs = (4096,) if quick else (1024, 4096, 16384, 65536)
    blocks = (64, 128, 256, 512, 1024)
    comps = (0,) if quick else (0, 1)
    lfs = (0.5,) if quick else (0.25, 0.5, 0.75)
    ref = None
    table = []
    for lf in lfs:
        batch.set_load_factor(lf)
        for V, blk, comp in itertools.product(Vs, blocks, comps):
            L = mgv.launch(0, blk, comp)
            g = batch.capture_fused(V, L, st.vp)
            ms = time_ms(lambda: mgv.call("mgv_graph_launch", g, st.vp), st)
            mgv.call("mgv_graph_destroy", g)
            r = batch.results()
            d = digest(r["count"], r["sum"])
            ref = check_same(ref, d, f"c4 candidate V={V} block={blk} compact={comp} lf={lf}")
            table.append({"vector_rows": V, "block": blk, "compact": comp, "load_factor": lf, "route_fanout": 1,
                          "ms": ms, "digest": d})
            log(f"c4 V={V} block={blk} compact={comp} lf={lf}: {ms:.3f} ms")
    best = min(table, key=lambda x: x["ms"])
    # the partition fan-out of the probes at the best launch: 1 = all fused
    batch.set_load_factor(best["load_factor"])
    L = mgv.launch(0, best["block"], best["compact"])
    thr, nst = 1 << 20, 4
    for fan in ((64,) if quick else (16, 64, 256)):
        g = batch.capture_routed(best["vector_rows"], L, st.vp, thr, nst, fan)
        ms = time_ms(lambda: mgv.call("mgv_graph_launch", g, st.vp), st)
        mgv.call("mgv_graph_destroy", g)
        r = batch.results()
        d = digest(r["count"], r["sum"])
        ref = check_same(ref, d, f"c4 routed fan-out {fan}")
        nbig = len(batch._routed(best["vector_rows"], thr, nst, fan)["big"])
        table.append({k: best[k] for k in ("vector_rows", "block", "compact", "load_factor")} |
                     {"route_fanout": fan, "route_threshold": thr, "route_streams": nst, "routed_queries": nbig,
                      "ms": ms, "digest": d})
        log(f"c4 routed fan-out {fan} ({nbig} queries): {ms:.3f} ms")
    # the LDS batch (builds and probes of at least min_rows rows through LDS
    # slices) at the best launch, per load factor, threshold and vector size
    # of the other (direct) queries: their few rows per query leave a fused
    # launch of large vectors latency-bound
    lds_table = []
    lds_Vs = sorted({1024, best["vector_rows"]})
    for lf in lfs:
        batch.set_load_factor(lf)
        for mr, lv in itertools.product((1 << 14,) if quick else (1 << 12, 1 << 14, 1 << 16), lds_Vs):
            g = batch.capture_ldsb(lv, L, st.vp, min_rows=mr)
            ms = time_ms(lambda: mgv.call("mgv_graph_launch", g, st.vp), st)
            mgv.call("mgv_graph_destroy", g)
            r = batch.results()
            d = digest(r["count"], r["sum"])
            ref = check_same(ref, d, f"c4 LDS batch lf={lf} min_rows={mr} V={lv}")
            lds_table.append({"load_factor": lf, "min_rows": mr, "vector_rows": lv, "ms": ms, "digest": d})
            log(f"c4 LDS batch lf={lf} min_rows={mr} V={lv}: {ms:.3f} ms")
    batch.close()
    best = min(table, key=lambda x: x["ms"])
    out = {k: best[k] for k in ("vector_rows", "block", "compact", "load_factor", "route_fanout")}
    if best["route_fanout"] > 1:
        out |= {"route_threshold": best["route_threshold"], "route_streams": best["route_streams"]}
    lb = min(lds_table, key=lambda x: x["ms"])
    out["lds_batch"] = {k: lb[k] for k in ("load_factor", "min_rows", "vector_rows", "ms")}
    out["path"] = "lds_batch" if lb["ms"] < best["ms"] else ("routed" if best["route_fanout"] > 1 else "fused")
    return out | {"ms": min(best["ms"], lb["ms"]), "queries": nq, "candidates": table + lds_table}


def tune_c2(st, quick=False, ns=1 << 24, nr=1 << 28):
    sk = mgv.DeviceBuffer(4 * ns)
    mgv.call("mgv_gen_perm", sk.vp, ns, ns, mgv.seed(2, 0), 0, st.vp)
    cols = {}
    for name, n, col, bits in (("sw", ns, 1, 16), ("rk", nr, 2, 0), ("rv", nr, 3, 16)):
        cols[name] = mgv.DeviceBuffer(4 * n)
        if name == "rk":
            mgv.call("mgv_gen_range", cols[name].vp, n, mgv.seed(2, col), 0, 2 * ns, st.vp)
        else:
            mgv.call("mgv_gen_bits", cols[name].vp, n, mgv.seed(2, col), 0, bits, st.vp)
    ht = mgv.HashTable(ns, 0.5)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, sk.vp, cols["sw"].vp, mgv.MGV_PAYLOAD_COLUMN, None, 0,
             ns, 0, res.vp, mgv.launch(0, 256), st.vp)
    st.sync()
    cands = []   # (mode, fan-out, block, coop, scatter shape, flags)
    for blk, coop in itertools.product((256,) if quick else (64, 256, 1024), (0, 1)):
        cands.append(("direct", 1, blk, coop, 0, 0))
    for f, shape in itertools.product((64,) if quick else (32, 64, 128, 256), (2,) if quick else (0, 1, 2, 3)):
        cands.append(("part", f, 256, 0, shape, 0))
    for f in ((32, 64) if quick else (16, 32, 64, 128)):
        cands.append(("lds", f, 1024, 0, 2, 0))
    cands.append(("ldsk", 32, 1024, 0, 2, MGV_LAUNCH_KEY_SLICES))
    for shape in ((2,) if quick else (1, 2, 3)):
        cands.append(("ldsc", 32, 1024, 0, shape, MGV_LAUNCH_COUNTED_PART))
    need = 0
    for mode, f, *_ in cands:
        b = C.c_uint64()
        if mode == "part":
            mgv.call("mgv_partition_workspace_size", nr, f, 2, C.byref(b))
        elif mode != "direct":
            mgv.call("mgv_probe_lds_workspace_size", nr, ht.nbuckets, f, C.byref(b))
        need = max(need, b.value)
    ws = mgv.DeviceBuffer(need)
    ref, table = None, []
    for mode, f, blk, coop, shape, flags in cands:
        L = mgv.Launch(0, blk, -1, coop | (shape << 8) | flags)

        def run():
            res.clear(st.vp)
            if mode == "direct":
                mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, None, 0, cols["rk"].vp, cols["rv"].vp, nr, 0,
                         0, res.vp, L, st.vp)
            elif mode == "part":
                mgv.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, None, 0, cols["rk"].vp,
                         cols["rv"].vp, nr, 0, 0, f, ws.vp, ws.nbytes, res.vp, L, st.vp)
            else:
                mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, None, 0, cols["rk"].vp, cols["rv"].vp,
                         nr, f, ws.vp, ws.nbytes, res.vp, L, st.vp)

        ms = time_ms(run, st)
        r = res.read()
        d = digest(np.array([r.count, r.sum], dtype=np.uint64))
        ref = check_same(ref, d, f"c2 candidate {mode} fanout={f} block={blk} coop={coop} scatter={shape}")
        table.append({"mode": mode, "fanout": f, "block": blk, "coop": coop, "scatter": shape, "ms": ms,
                      "digest": d})
        log(f"c2 {mode} fanout={f} block={blk} coop={coop} scatter={shape}: {ms:.3f} ms")
    for b in (sk, ht.buf, res, ws, *cols.values()):
        b.close()
    best = min(table, key=lambda x: x["ms"])
    return {k: best[k] for k in ("mode", "fanout", "block", "coop", "scatter")} | {"ms": best["ms"],
                                                                                  "candidates": table}


def tune_c1(st, quick=False, n=1 << 28):
    a = mgv.DeviceBuffer(4 * n)
    b = mgv.DeviceBuffer(4 * n)
    mgv.call("mgv_gen_bits", a.vp, n, mgv.seed(1, 0), 0, 20, st.vp)
    mgv.call("mgv_gen_bits", b.vp, n, mgv.seed(1, 1), 0, 31, st.vp)
    res = mgv.ResultBuffer()
    ref, table = None, []
    for blk, comp in itertools.product((64, 256, 1024), (0, 1)):
        L = mgv.launch(0, blk, comp)

        def run():
            res.clear(st.vp)
            mgv.call("mgv_select_sum", b.vp, a.vp, n, (1 << 30) - 1, res.vp, L, st.vp)

        ms = time_ms(run, st)
        r = res.read()
        d = digest(np.array([r.count, r.sum], dtype=np.uint64))
        ref = check_same(ref, d, f"c1 candidate block={blk} compact={comp}")
        table.append({"block": blk, "compact": comp, "ms": ms, "digest": d})
        log(f"c1 block={blk} compact={comp}: {ms:.4f} ms")
    for x in (a, b, res):
        x.close()
    best = min(table, key=lambda x: x["ms"])
    return {k: best[k] for k in ("block", "compact")} | {"ms": best["ms"], "candidates": table}


TUNERS = {"c4": tune_c4, "c2": tune_c2, "c1": tune_c1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="c4,c2,c1")
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--quick", action="store_true", help="reduced search spaces (same workloads)")
    args = ap.parse_args()
    mgv.load()
    mgv.load_ops()
    mgv.init(0)
    st = mgv.Stream()
    try:
        with open(args.out) as f:
            out = json.load(f)
    except (OSError, ValueError):
        out = {}
    p = mgv.props()
    for w in [x for x in args.workloads.split(",") if x]:
        t0 = time.time()
        out[w] = TUNERS[w](st, args.quick)
        out[w]["device"] = p.arch.decode()
        out[w]["tuned_s"] = round(time.time() - t0, 1)
        log(f"{w}: best {({k: v for k, v in out[w].items() if k != 'candidates'})}")
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({w: {k: v for k, v in out[w].items() if k != "candidates"} for w in out}))


if __name__ == "__main__":
    main()
