"""Autotuner of the launch parameters (run on the GPU; results persisted to
tuned.json beside this file, which bench.py reads).

Workloads and search spaces:
  c4  256-query select -> join -> SUM batch in one fused graph:
      vector size {1K, 4K, 16K, 64K} x workgroup {64, 256, 1024} x
      select compaction {0, 1} x hash-table load factor {0.25, 0.5, 0.75}
  c2  2^28-row probe of a 2^24-key table (load factor 0.5): the direct probe
      (workgroup {64, 256, 1024} x {per-lane, 4-lane cooperative}), the
      partitioned probe (fan-out {32..256} x scatter shape) and the two-level
      LDS probe (fan-out {16..128} x scatter shape)
  c1  filtered-sum scan of 2^28 rows: workgroup x compaction

Every candidate must give results identical to the first candidate's
(bit-exact COUNT / SUM per query); a candidate that disagrees is an error,
not a slower entry.

  python multigpu-vector_amd/autotune.py --workloads c4,c2,c1
"""
from __future__ import annotations

import argparse
import ctypes as C
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


def log(*a):
    print("[autotune]", *a, file=sys.stderr, flush=True)


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


def tune_c4(st, quick=False):
    shapes = workloads.c4_shapes(256)
    batch = workloads.JoinBatch(shapes, 0.5, st.vp)
    st.sync()
    Vs = (4096, 16384) if quick else (1024, 4096, 16384, 65536)
    blocks = (256,) if quick else (64, 256, 1024)
    lfs = (0.5,) if quick else (0.25, 0.5, 0.75)
    ref = None
    table = []
    for lf in lfs:
        batch.set_load_factor(lf)
        for V, blk, comp in itertools.product(Vs, blocks, (0, 1)):
            L = mgv.launch(0, blk, comp)
            g = batch.capture_fused(V, L, st.vp)
            ms = time_ms(lambda: mgv.call("mgv_graph_launch", g, st.vp), st)
            mgv.call("mgv_graph_destroy", g)
            r = batch.results()
            key = (r["count"].tobytes(), r["sum"].tobytes())
            if ref is None:
                ref = key
            elif key != ref:
                raise RuntimeError(f"c4 candidate V={V} block={blk} compact={comp} lf={lf} changed the results")
            table.append({"vector_rows": V, "block": blk, "compact": comp, "load_factor": lf, "ms": ms})
            log(f"c4 V={V} block={blk} compact={comp} lf={lf}: {ms:.3f} ms")
    batch.close()
    best = min(table, key=lambda x: x["ms"])
    return {k: best[k] for k in ("vector_rows", "block", "compact", "load_factor")} | {"ms": best["ms"],
                                                                                       "candidates": table}


def tune_c2(st, quick=False):
    ns, nr = 1 << 24, 1 << 28
    sk = mgv.DeviceBuffer(4 * ns)
    mgv.call("mgv_gen_perm", sk.vp, ns, ns, mgv.seed(2, 0), 0, st.vp)
    cols = {}
    for name, n, col, bits in (("sw", ns, 1, 16), ("rk", nr, 2, 25), ("rv", nr, 3, 16)):
        cols[name] = mgv.DeviceBuffer(4 * n)
        mgv.call("mgv_gen_bits", cols[name].vp, n, mgv.seed(2, col), 0, bits, st.vp)
    ht = mgv.HashTable(ns, 0.5)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, sk.vp, cols["sw"].vp, mgv.MGV_PAYLOAD_COLUMN, None, 0,
             ns, 0, res.vp, mgv.launch(0, 256), st.vp)
    st.sync()
    cands = []
    for blk, coop in itertools.product((256,) if quick else (64, 256, 1024), (0, 1)):
        cands.append(("direct", 1, blk, coop, 0))
    for f, shape in itertools.product((64,) if quick else (32, 64, 128, 256), (0, 2) if quick else (0, 1, 2, 3)):
        cands.append(("part", f, 256, 0, shape))
    for f, shape in itertools.product((32, 64) if quick else (16, 32, 64, 128), (0, 2) if quick else (0, 1, 2, 3)):
        cands.append(("lds", f, 1024, 0, shape))
    need = 0
    for mode, f, _, _, _ in cands:
        b = C.c_uint64()
        if mode == "part":
            mgv.call("mgv_partition_workspace_size", nr, f, 2, C.byref(b))
        elif mode == "lds":
            mgv.call("mgv_probe_lds_workspace_size", nr, ht.nbuckets, f, C.byref(b))
        need = max(need, b.value)
    ws = mgv.DeviceBuffer(need)
    ref, table = None, []
    for mode, f, blk, coop, shape in cands:
        L = mgv.Launch(0, blk, -1, coop | (shape << 8))

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
        if ref is None:
            ref = (r.count, r.sum)
        elif (r.count, r.sum) != ref:
            raise RuntimeError(f"c2 candidate {mode} fanout={f} block={blk} coop={coop} changed the result")
        table.append({"mode": mode, "fanout": f, "block": blk, "coop": coop, "scatter": shape, "ms": ms})
        log(f"c2 {mode} fanout={f} block={blk} coop={coop} scatter={shape}: {ms:.3f} ms")
    for b in (sk, ht.buf, res, ws, *cols.values()):
        b.close()
    best = min(table, key=lambda x: x["ms"])
    return {k: best[k] for k in ("mode", "fanout", "block", "coop", "scatter")} | {"ms": best["ms"],
                                                                                  "candidates": table}


def tune_c1(st, quick=False):
    n = 1 << 28
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
        if ref is None:
            ref = (r.count, r.sum)
        elif (r.count, r.sum) != ref:
            raise RuntimeError("c1 candidate changed the result")
        table.append({"block": blk, "compact": comp, "ms": ms})
        log(f"c1 block={blk} compact={comp}: {ms:.4f} ms")
    for x in (a, b, res):
        x.close()
    best = min(table, key=lambda x: x["ms"])
    return {k: best[k] for k in ("block", "compact")} | {"ms": best["ms"], "candidates": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="c4,c2,c1")
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--quick", action="store_true")
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
        out[w] = {"c4": tune_c4, "c2": tune_c2, "c1": tune_c1}[w](st, args.quick)
        out[w]["device"] = p.arch.decode()
        out[w]["tuned_s"] = round(time.time() - t0, 1)
        log(f"{w}: best {({k: v for k, v in out[w].items() if k != 'candidates'})}")
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({w: {k: v for k, v in out[w].items() if k != "candidates"} for w in out}))


if __name__ == "__main__":
    main()
