#!/usr/bin/env python3
"""Benchmark of the select -> hash-join -> aggregate path on one MI355X.

Headline (one JSON line on stdout): probe throughput of the C2 workload --
SELECT COUNT(*), SUM(R.v + S.w) FROM R JOIN S ON R.k = S.k with |S| = 2^24
(keys a seeded permutation of 0..2^24-1, w in [0,2^16)), |R| = 2^28 (k in
[0,2^25), v in [0,2^16)), open addressing at load factor 0.5.  A step is one
probe pass of R against the resident table (inputs resident in HBM), timed K
times between device synchronisations; the build is timed separately.

Extra legs (outside the timed steps) report the filtered-sum scan (C1) and
later configurations as extra JSON keys.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multigpu-vector_amd"))

import numpy as np  # noqa: E402

import mgv  # noqa: E402

HBM_PEAK_GBPS = 8000.0
MGV_LAUNCH_COUNTED_PART = 1 << 12
MGV_LAUNCH_KEY_SLICES = 1 << 13


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def gb(x):
    return x / 1e9


class Timer:
    """HIP events on one stream: per-launch kernel durations."""

    def __init__(self, stream):
        self.s = stream
        self.ev = []

    def around(self, fn):
        a, b = mgv.Event(), mgv.Event()
        a.record(self.s)
        fn()
        b.record(self.s)
        self.ev.append((a, b))

    def durations_ms(self):
        out = [a.elapsed_ms(b) for a, b in self.ev]
        for a, b in self.ev:
            a.close()
            b.close()
        self.ev = []
        return out


def gen_bits(n, seed, stream, bits, s=None):
    buf = mgv.DeviceBuffer(4 * n)
    mgv.call("mgv_gen_bits", buf.vp, n, seed, stream, bits, s)
    return buf


def gen_perm(n, rng, seed, stream, s=None):
    buf = mgv.DeviceBuffer(4 * n)
    mgv.call("mgv_gen_perm", buf.vp, n, rng, seed, stream, s)
    return buf


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/traffic.json)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


# ------------------------------------------------------------------ C2 join

def leg_c2(args, st):
    ns, nr = 1 << 24, 1 << 28
    log(f"C2: generating S ({ns} rows) and R ({nr} rows)")
    sk = gen_perm(ns, ns, mgv.seed(2, 0), 0, st.vp)
    sw = gen_bits(ns, mgv.seed(2, 1), 0, 16, st.vp)
    rk = gen_bits(nr, mgv.seed(2, 2), 0, 25, st.vp)
    rv = gen_bits(nr, mgv.seed(2, 3), 0, 16, st.vp)
    ht = mgv.HashTable(ns, args.load_factor)
    res = mgv.ResultBuffer()
    st.sync()
    t2 = tuned("c2") or {}
    block = args.block or int(t2.get("block", 256))
    coop = int(t2.get("coop", 0)) if args.coop < 0 else args.coop
    shape = int(t2.get("scatter", 2)) if args.scatter < 0 else args.scatter
    L = mgv.Launch(0, block, -1, coop | (shape << 8))
    Llds = mgv.Launch(0, 1024, -1, shape << 8)

    # build (timed: clear excluded): the direct build and the build through
    # LDS table slices; the faster one's table is kept for the probe
    bneed = C.c_uint64()
    mgv.call("mgv_build_lds_workspace_size", ns, ht.nbuckets, 32, C.byref(bneed))
    bws = mgv.DeviceBuffer(bneed.value)
    Lb = mgv.Launch(0, 1024, -1, 2 << 8)
    builds = {
        "direct": lambda: mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, sk.vp, sw.vp,
                                   mgv.MGV_PAYLOAD_COLUMN, None, 0, ns, 0, res.vp, L, st.vp),
        "lds": lambda: mgv.call("mgv_hashjoin_build_lds", ht.buf.vp, ht.nbuckets, sk.vp, sw.vp,
                                mgv.MGV_PAYLOAD_COLUMN, None, 0, ns, 0, 32, bws.vp, bws.nbytes, res.vp, Lb, st.vp),
    }
    build_by = {}
    for name, fn in builds.items():
        bms = []
        for rep in range(3):
            ht.clear(st.vp)
            res.clear(st.vp)
            t = Timer(st.vp)
            t.around(fn)
            st.sync()
            bms.append(t.durations_ms()[0])
        assert res.read().errors == 0
        build_by[name] = min(bms[1:])
        log(f"C2: build {name} {build_by[name]:.3f} ms ({gb(ns) / (build_by[name] / 1e3):.2f} Grows/s)")
    build_path = min(build_by, key=build_by.get)
    build_t = build_by[build_path]
    ht.clear(st.vp)
    res.clear(st.vp)
    builds[build_path]()
    st.sync()
    assert res.read().errors == 0
    bws.close()

    ws = None

    def make_probe(mode, fanout):
        """mode "direct" (one pass over R), "part" (one-level partitioned,
        fan-out F), "lds" (two-level count-free partitioning, slices probed
        from LDS), "ldsc" (the same with counted partitionings) or "ldsk"
        (count-free, slices held as 32-bit key hashes instead of slot tags)."""
        nonlocal ws
        need = C.c_uint64()
        if mode == "part":
            mgv.call("mgv_partition_workspace_size", nr, fanout, 2, C.byref(need))
        elif mode in ("lds", "ldsc", "ldsk"):
            mgv.call("mgv_probe_lds_workspace_size", nr, ht.nbuckets, fanout, C.byref(need))
        if need.value and (ws is None or ws.nbytes < need.value):
            if ws is not None:
                ws.close()
            ws = mgv.DeviceBuffer(need.value)
        if mode == "part":
            return lambda: mgv.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp,
                                    nr, 0, 0, fanout, ws.vp, ws.nbytes, res.vp, L, st.vp)
        if mode in ("lds", "ldsc", "ldsk"):
            flag = {"lds": 0, "ldsc": MGV_LAUNCH_COUNTED_PART, "ldsk": MGV_LAUNCH_KEY_SLICES}[mode]
            Lm = mgv.Launch(0, 1024, -1, Llds.reserved | flag)
            return lambda: mgv.call("mgv_hashjoin_probe_lds", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp, nr,
                                    fanout, ws.vp, ws.nbytes, res.vp, Lm, st.vp)
        return lambda: mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp, nr, 0, 0,
                                res.vp, L, st.vp)

    # choose the probe path
    if args.probe != "auto":
        candidates = [(args.probe, args.fanout if args.fanout > 0 else (1 if args.probe == "direct" else 64))]
    else:
        candidates = [("direct", 1), ("part", 64), ("ldsc", 32), ("ldsk", 32), ("lds", 32), ("lds", 64)]
        if "mode" in t2 and (t2["mode"], int(t2["fanout"])) not in candidates:
            candidates.append((t2["mode"], int(t2["fanout"])))
    sweep, sweep_res = {}, {}
    for mode, f in candidates:
        fn = make_probe(mode, f)
        t = Timer(st.vp)
        for rep in range(3):
            res.clear(st.vp)
            t.around(fn)
        st.sync()
        sweep[(mode, f)] = min(t.durations_ms()[1:] or t.durations_ms())
        rr = res.read()
        sweep_res[(mode, f)] = (rr.count, rr.sum)
        log(f"C2: probe {mode} fan-out {f}: {sweep[(mode, f)]:.3f} ms, count {rr.count} sum {rr.sum}")
    best_mode, fanout = min(sweep, key=sweep.get)
    probe_fn = make_probe(best_mode, fanout)

    def probe_step(timer=None):
        res.clear(st.vp)
        if timer:
            timer.around(probe_fn)
        else:
            probe_fn()

    for _ in range(args.warmup):
        probe_step()
    st.sync()
    # timed steps: wall clock around K steps, events around each probe call
    timer = Timer(st.vp)
    mgv.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        probe_step(timer)
    st.sync()
    mgv.sync()
    wall = time.perf_counter() - t0
    kms = timer.durations_ms()
    r = res.read()

    # parity (checker, on the CPU): the closed form of the C2 join, COUNT =
    # #{R.k < 2^24} and SUM = sum of R.v + w[perm^-1(R.k)] over those rows,
    # from the inputs the GPU used; the timed steps' result and every swept
    # candidate's must equal it
    cf = closed_form_c2(args, st, sk, sw, rk, rv, ns, nr)
    log(f"C2: probe count {r.count} sum {r.sum}; closed form count {cf[0]} sum {cf[1]}")
    parity = {
        "closed_form_count": cf[0], "closed_form_sum": cf[1],
        "timed_path_equals_closed_form": (r.count, r.sum) == cf,
        "every_candidate_equals_closed_form": all(v == cf for v in sweep_res.values()),
        "candidates": {f"{m}{f}": list(v) for (m, f), v in sweep_res.items()},
    }

    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_ms = float(np.mean(kms))
    algo_bytes = nr * 16
    out = {
        "ms_per_step": ms_per_step,
        "value": nr / (ms_per_step / 1e3) / 1e9,
        "kernel_ms_avg": avg_kernel_ms,
        "kernel_ms_min": float(np.min(kms)),
        "kernel_ms_median": float(np.median(kms)),
        "roofline": {
            "bound": "hbm",
            "achieved": algo_bytes / (avg_kernel_ms / 1e3) / 1e9,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": algo_bytes / (avg_kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBPS,
            "traffic": traffic_from_profiles(f"probe_c2_{best_mode}{fanout}"),
            # the measured HBM traffic per call over the same average call time
            "traffic_gbps": (traffic_from_profiles(f"probe_c2_{best_mode}{fanout}") or 0) / (avg_kernel_ms / 1e3) / 1e9
            or None,
            "traffic_frac": (traffic_from_profiles(f"probe_c2_{best_mode}{fanout}") or 0) / (avg_kernel_ms / 1e3) / 1e9
            / HBM_PEAK_GBPS or None,
            "kernel": {"direct": "k_pipeline<PProbeSum<false,true,false>> (mgv_hashjoin_probe)",
                       "part": "mgv_hashjoin_probe_partitioned: k_part_prep + k_part_count + k_part_scan + "
                               "k_part_scatter + k_part_probe (events around the whole call)",
                       "ldsc": "mgv_hashjoin_probe_lds (counted): two partitionings (k_part_prep/count/scan/"
                               "scatter x 2) + k_lds_probe (events around the whole call)",
                       "lds": "mgv_hashjoin_probe_lds: counter fill + two count-free scatters (k_fill_scatter x 2) + "
                              "k_lds_probe + k_probe_hashed (events around the whole call)"}[best_mode],
            "probe_path": best_mode,
            "fanout": fanout,
            "scatter_shape": shape,
            "block": block,
            "coop_probe": coop,
            "algorithmic_bytes_per_launch": algo_bytes,
        },
        "build": {"ms": build_t, "path": build_path, "ms_by_path": build_by,
                  "grows_per_s": gb(ns) / (build_t / 1e3),
                  "gbps_algorithmic": (ns * 8 + ht.nbuckets * 64) / (build_t / 1e3) / 1e9},
        "result": {"count": r.count, "sum": r.sum, **parity},
        "probe_path_sweep_ms": {f"{m}{f}": v for (m, f), v in sweep.items()},
    }

    if args.cpu:
        out["cpu_c2"] = cpu_baseline_c2(args, st, sk, sw, rk, rv, ht, ns, nr)
        out["cpu_c2"]["equals_gpu"] = out["cpu_c2"]["result"] == [r.count, r.sum]
    for b in (sk, sw, rk, rv, ht.buf, res) + ((ws,) if ws is not None else ()):
        b.close()
    return out


def cpu_threads(args):
    """Threads of the all-threads CPU figures: the CPUs this process may run
    on, limited to the host share given to one GPU (OMP_NUM_THREADS, which
    the GPU host sets to its per-GPU share)."""
    if args.cpu_threads:
        return args.cpu_threads
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_info():
    info = {"hardware_concurrency": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        keys = {"model name": "cpu_model", "cpu family": "cpu_family", "model": "cpu_model_id",
                "stepping": "stepping", "cpu MHz": "cpu_mhz"}
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in keys and keys[k] not in info:
                info[keys[k]] = v.strip()
            if not line.strip() and "cpu_model" in info:
                break
    except OSError:
        pass
    return info


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    return oracle_py.Oracle()


def closed_form_c2(args, st, sk, sw, rk, rv, ns, nr):
    """Checker: (COUNT, SUM) of the C2 join by direct addressing on the CPU
    (the build keys are a permutation of [0, ns))."""
    st.sync()
    hk, hw = sk.download(np.int32, ns), sw.download(np.int32, ns)
    hrk, hrv = rk.download(np.int32, nr), rv.download(np.int32, nr)
    got = oracle().join_dense(hk, hw, ns, hrk, hrv, threads=cpu_threads(args))
    assert got is not None, "C2 build keys are not a permutation"
    return got


def cpu_baseline_c2(args, st, sk, sw, rk, rv, ht, ns, nr):
    """CPU baseline of the C2 join: the oracle's chained hash table (the
    reference's own CPU algorithm, main8.cu:84-103, and its device build,
    hashjoin.cu:41-60).  All threads: a threaded build of all S rows and a
    probe of all R rows.  One thread: the build and a probe of the first
    cpu_sample R rows.  Both must give the GPU's results on the same rows."""
    orc = oracle()
    st.sync()
    hk, hw = sk.download(np.int32, ns), sw.download(np.int32, ns)
    hrk, hrv = rk.download(np.int32, nr), rv.download(np.int32, nr)
    T = cpu_threads(args)
    t0 = time.perf_counter()
    t = orc.table(hk, hw, log_buckets=25, threads=T)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    want = orc.probe(t, hrk, hrv, threads=T)
    t_probe = time.perf_counter() - t0
    orc.free(t)
    M = min(args.cpu_sample, nr)
    t0 = time.perf_counter()
    t1 = orc.table(hk, hw, log_buckets=25)
    t_build1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    want1 = orc.probe(t1, hrk[:M], hrv[:M])
    t_probe1 = time.perf_counter() - t0
    orc.free(t1)
    res = mgv.ResultBuffer()
    mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, None, 0, rk.vp, rv.vp, M, 0, 0, res.vp,
             mgv.launch(0, 256), st.vp)
    st.sync()
    g = res.read()
    res.close()
    log(f"C2 cpu: {T} threads build {t_build:.2f} s probe {nr} rows {t_probe:.2f} s; 1 thread build "
        f"{t_build1:.2f} s probe {M} rows {t_probe1:.2f} s")
    return {
        "threads": T, "build_s": t_build, "probe_s": t_probe, "probe_rows": nr,
        "grows_per_s": nr / t_probe / 1e9,
        "one_thread": {"build_s": t_build1, "probe_s": t_probe1, "probe_rows": M, "grows_per_s": M / t_probe1 / 1e9},
        "result": [want["count"], want["sum"]],
        "one_thread_sample_equals_gpu": (g.count, g.sum) == (want1["count"], want1["sum"]),
    }


def cpu_baseline_c1(args, a, b, n, bound, gpu):
    """CPU baseline of C1, exactly the configuration: SELECT SUM(a) WHERE b <=
    2^30 - 1 over 2^24 rows, int64 accumulator (the oracle's sum_select_cpu,
    main9.cu:56-61), on one thread and on all threads; best of 5 runs each."""
    orc = oracle()
    ha, hb = a.download(np.int32, n), b.download(np.int32, n)
    T = cpu_threads(args)
    out = {"rows": n, "threads": T}
    for name, fn in (("one_thread", lambda: orc.select_sum(hb, ha, bound)[:2]),
                     ("all_threads", lambda: orc.select_sum_mt(hb, ha, bound, T))):
        ts, got = [], None
        for _ in range(5):
            t0 = time.perf_counter()
            got = fn()
            ts.append(time.perf_counter() - t0)
        ms = min(ts) * 1e3
        out[name] = {"ms": ms, "grows_per_s": n / (ms / 1e3) / 1e9, "gbps": n * 8 / (ms / 1e3) / 1e9,
                     "count": int(got[0]), "sum": int(got[1]), "equals_gpu": (int(got[0]), int(got[1])) == gpu}
    log(f"C1 cpu: 1 thread {out['one_thread']['ms']:.2f} ms, {T} threads {out['all_threads']['ms']:.2f} ms "
        f"for {n} rows; equal to GPU {out['one_thread']['equals_gpu'] and out['all_threads']['equals_gpu']}")
    return out


# ------------------------------------------------------------------ C1 scan

def leg_c1(args, st):
    out = {}
    bound = (1 << 30) - 1
    t1 = tuned("c1") or {}
    block = int(t1.get("block", 256))
    for n, reps_in_graph in ((1 << 24, 100), (1 << 28, 1)):
        a = gen_bits(n, mgv.seed(1, 0), 0, 20, st.vp)
        b = gen_bits(n, mgv.seed(1, 1), 0, 31, st.vp)
        res = mgv.ResultBuffer(reps_in_graph)
        st.sync()
        best = None
        for compact in (0, 1):
            L = mgv.launch(0, block, compact)
            # capture reps_in_graph back-to-back scans in one graph
            mgv.call("mgv_graph_begin_capture", st.vp)
            for i in range(reps_in_graph):
                mgv.call("mgv_select_sum", b.vp, a.vp, n, bound, res.slot(i), L, st.vp)
            g = C.c_void_p()
            mgv.call("mgv_graph_end_capture", st.vp, C.byref(g))
            times = []
            for rep in range(4):
                res.clear(st.vp)
                t = Timer(st.vp)
                t.around(lambda: mgv.call("mgv_graph_launch", g, st.vp))
                st.sync()
                times.append(t.durations_ms()[0] / reps_in_graph)
            mgv.call("mgv_graph_destroy", g)
            r = res.read(0)
            kms = min(times[1:])
            gbps = n * 8 / (kms / 1e3) / 1e9
            log(f"C1 n={n} compact={compact}: {kms:.4f} ms/scan, {gbps:.0f} GB/s, count {r.count} sum {r.sum}")
            if best is None or kms < best["ms"]:
                best = {"ms": kms, "gbps": gbps, "frac": gbps / HBM_PEAK_GBPS, "compact": compact, "block": block,
                        "count": r.count, "sum": r.sum, "grows_per_s": n / (kms / 1e3) / 1e9}
        key = "c1_2^24_graph100" if n == 1 << 24 else "c1_2^28"
        out[key] = best
        if n == 1 << 24 and args.cpu:
            cpu = cpu_baseline_c1(args, a, b, n, bound, (best["count"], best["sum"]))
            out["cpu"] = cpu
            best["vs_cpu_all_threads"] = best["grows_per_s"] / cpu["all_threads"]["grows_per_s"]
            best["vs_cpu_one_thread"] = best["grows_per_s"] / cpu["one_thread"]["grows_per_s"]
        for x in (a, b, res):
            x.close()
    out["c1_fused_batch"] = c1_fused_batch(args, st, bound)
    return out


def c1_fused_batch(args, st, bound):
    """The filtered sum through the fused batch kernel: 64 independent
    SELECT COUNT(*), SUM(a) WHERE b <= bound queries of 2^22 rows each (2^28
    rows, 2 GiB) drained by one persistent mgv_batch_run launch in a graph;
    the results must equal 64 separate mgv_select_sum calls."""
    nq, rows = 64, 1 << 22
    n = nq * rows
    a = gen_bits(n, mgv.seed(1, 0), 0, 20, st.vp)
    b = gen_bits(n, mgv.seed(1, 1), 0, 31, st.vp)
    res = mgv.ResultBuffer(2 * nq)
    q = (mgv.Query * nq)()
    for i in range(nq):
        q[i] = mgv.Query(mgv.MGV_Q_SELECT_SUM, bound, b.ptr + 4 * i * rows, None, a.ptr + 4 * i * rows, rows, None, 0,
                         0, None, None, res.ptr + 64 * i, 0, 0, 0)
    best = None
    for V in (65536, 262144):
        cnt = C.c_uint64()
        mgv.call("mgv_batch_plan", q, nq, V, None, 0, C.byref(cnt))
        items = (mgv.WorkItem * cnt.value)()
        mgv.call("mgv_batch_plan", q, nq, V, items, cnt.value, C.byref(cnt))
        dq = mgv.DeviceBuffer(C.sizeof(q)).upload(np.frombuffer(bytes(q), dtype=np.uint8))
        di = mgv.DeviceBuffer(C.sizeof(items)).upload(np.frombuffer(bytes(items), dtype=np.uint8))
        ctr = mgv.DeviceBuffer(256)
        for blk in (256, 1024):
            L = mgv.launch(0, blk, 0)
            mgv.call("mgv_graph_begin_capture", st.vp)
            mgv.call("mgv_result_clear", res.vp, nq, st.vp)
            mgv.call("mgv_batch_run", dq.vp, di.vp, cnt.value, ctr.vp, L, st.vp)
            g = C.c_void_p()
            mgv.call("mgv_graph_end_capture", st.vp, C.byref(g))
            times = []
            for rep in range(4):
                t = Timer(st.vp)
                t.around(lambda: mgv.call("mgv_graph_launch", g, st.vp))
                st.sync()
                times.append(t.durations_ms()[0])
            mgv.call("mgv_graph_destroy", g)
            ms = min(times[1:])
            if best is None or ms < best["ms"]:
                best = {"ms": ms, "vector_rows": V, "block": blk}
        for x in (dq, di, ctr):
            x.close()
    fused = [(res.read(i).count, res.read(i).sum) for i in range(nq)]
    for i in range(nq):
        mgv.call("mgv_select_sum", b.at(4 * i * rows), a.at(4 * i * rows), rows, bound, res.slot(nq + i),
                 mgv.launch(0, 1024, 0), st.vp)
    st.sync()
    plain = [(res.read(nq + i).count, res.read(nq + i).sum) for i in range(nq)]
    gbps = n * 8 / (best["ms"] / 1e3) / 1e9
    best.update({"queries": nq, "rows_per_query": rows, "gbps": gbps, "frac": gbps / HBM_PEAK_GBPS,
                 "results_equal_per_query_calls": fused == plain})
    log(f"C1 fused batch of {nq} x 2^22: {best['ms']:.4f} ms, {gbps:.0f} GB/s ({gbps / HBM_PEAK_GBPS:.0%}), "
        f"equal {fused == plain}")
    for x in (a, b, res):
        x.close()
    return best


# ------------------------------------------------------------------ host-resident columns

def leg_f1(args, st):
    """The reference's path for host-resident inputs (exchange.cu:196-291):
    generator -> exchange -> consumer (double-buffered pinned H2D on a copy
    stream, events) -> select -> aggregation -> gpu_to_cpu, over 2^26 rows of
    {b, a} in pinned host memory, vectors of 2 Mi rows (two columns fill one
    4 Mi-row device buffer of the consumer's staging; mgv_ops_sum_select).
    Reported: its rate; the same plan over device-resident columns; a plain
    H2D copy of the same bytes; and the overlap factor (copy time + device
    time) / host-column time, above 1 when copies and kernels overlap."""
    n, V = 1 << 26, 2 << 20
    mgv.ops_call("mgv_ops_init", 8, 2, 4)
    a = gen_bits(n, mgv.seed(1, 0), 7, 20, st.vp)
    b = gen_bits(n, mgv.seed(1, 1), 7, 31, st.vp)
    st.sync()
    ha, hb = C.c_void_p(), C.c_void_p()
    mgv.call("mgv_host_malloc", C.byref(ha), 4 * n)
    mgv.call("mgv_host_malloc", C.byref(hb), 4 * n)
    mgv.call("mgv_memcpy", ha, a.vp, 4 * n)
    mgv.call("mgv_memcpy", hb, b.vp, 4 * n)
    bound = (1 << 30) - 1
    dst = np.zeros(4, np.int32)

    def plan(pb, pa, where):
        s = C.c_int64()
        t0 = time.perf_counter()
        mgv.ops_call("mgv_ops_sum_select", pb, pa, n, bound, V, where, 256, 1024, dst.ctypes.data_as(C.c_void_p),
                     C.byref(s))
        return time.perf_counter() - t0, s.value

    host = [plan(hb, ha, -1) for _ in range(3)]
    dev = [plan(b.vp, a.vp, 0) for _ in range(3)]
    copy = []
    for _ in range(3):
        t0 = time.perf_counter()
        mgv.call("mgv_memcpy", b.vp, hb, 4 * n)
        mgv.call("mgv_memcpy", a.vp, ha, 4 * n)
        copy.append(time.perf_counter() - t0)
    th, tdv, tc = min(x[0] for x in host[1:]), min(x[0] for x in dev[1:]), min(copy[1:])
    sums = {x[1] for x in host + dev}
    out = {"rows": n, "vector_rows": V, "bytes": 8 * n,
           "host_columns": {"ms": th * 1e3, "grows_per_s": n / th / 1e9, "h2d_gbps": 8 * n / th / 1e9},
           "device_columns": {"ms": tdv * 1e3, "grows_per_s": n / tdv / 1e9},
           "plain_h2d_copy": {"ms": tc * 1e3, "gbps": 8 * n / tc / 1e9},
           "overlap": (tc + tdv) / th, "sums_equal": len(sums) == 1}
    log(f"F1 host columns {th * 1e3:.1f} ms ({8 * n / th / 1e9:.1f} GB/s), device columns {tdv * 1e3:.1f} ms, "
        f"plain copy {tc * 1e3:.1f} ms ({8 * n / tc / 1e9:.1f} GB/s), overlap {out['overlap']:.2f}, "
        f"sums equal {out['sums_equal']}")
    mgv.call("mgv_host_free", ha)
    mgv.call("mgv_host_free", hb)
    a.close()
    b.close()
    return out


# ------------------------------------------------------------------ copy

def leg_copy(args, st):
    """Plain device-to-device copies (mgv_memcpy_async) of 1, 2 and 4 GiB:
    the read+write bandwidth that a one-pass scatter can reach (each of the
    probe's partitionings reads and writes its rows once), and the
    calibration of rocprofv3's FETCH_SIZE / WRITE_SIZE counters (a copy of B
    bytes reads B and writes B)."""
    out = {}
    for nbytes in (1 << 30, 2 << 30, 4 << 30):
        a, b = mgv.DeviceBuffer(nbytes), mgv.DeviceBuffer(nbytes)
        mgv.call("mgv_fill_u32", a.vp, nbytes // 4, 7, st.vp)
        t = Timer(st.vp)
        for _ in range(6):
            t.around(lambda: mgv.call("mgv_memcpy_async", b.vp, a.vp, nbytes, st.vp))
        st.sync()
        ms = min(t.durations_ms()[1:])
        out[f"{nbytes >> 30}GiB"] = {"ms": ms, "gbps_read_write": 2 * nbytes / ms / 1e6}
        a.close()
        b.close()
    log("copy: " + ", ".join(f"{k} {v['ms']:.3f} ms ({v['gbps_read_write']:.0f} GB/s read+write)"
                             for k, v in out.items()))
    return out


# ------------------------------------------------------------------ split

def leg_f3(args, st):
    """The split operator (mgv_split): 2^28 rows routed to two parents by
    parity, (a) one column and no filter, 8 B per row (4 read, 4 written);
    (b) pred <= 2^30-1 then the value column, selectivity 1/2, 10 B per row
    (8 read, 2 written on average); (c) two carried columns, no filter,
    16 B per row.  Cursors are cleared outside the timed
    launch; the kept-row total is checked against n and mgv_select_sum."""
    out = {}
    n = 1 << 28
    bound = (1 << 30) - 1
    x = gen_bits(n, mgv.seed(6, 0), 0, 32, st.vp)
    pred = gen_bits(n, mgv.seed(6, 1), 0, 31, st.vp)
    outs = [mgv.DeviceBuffer(4 * n) for _ in range(2)]
    cnt = mgv.DeviceBuffer(64)
    res = mgv.ResultBuffer(1)
    mgv.call("mgv_select_sum", pred.vp, x.vp, n, bound, res.vp, mgv.launch(0, 1024, 0), st.vp)
    st.sync()
    selected = res.read(0).count
    # one column; or (x, pred) carried as two columns into four outputs
    cp1 = (C.c_void_p * 1)(x.ptr)
    op1 = (C.c_void_p * 2)(outs[0].ptr, outs[1].ptr)
    extra_outs = [mgv.DeviceBuffer(4 * n) for _ in range(2)]
    cp2 = (C.c_void_p * 2)(x.ptr, pred.ptr)
    op2 = (C.c_void_p * 4)(outs[0].ptr, extra_outs[0].ptr, outs[1].ptr, extra_outs[1].ptr)
    for name, p, ncols, bytes_per_row, want in (("split_2^28", None, 1, 8.0, n),
                                                ("split_2^28_filtered", pred, 1, 10.0, selected),
                                                ("split_2^28_2cols", None, 2, 16.0, n)):
        cp, op = (cp1, op1) if ncols == 1 else (cp2, op2)
        times = []
        for rep in range(5):
            mgv.call("mgv_fill_u64", cnt.vp, 2, 0, st.vp)
            t = Timer(st.vp)
            t.around(lambda: mgv.call("mgv_split", p.vp if p else None, bound, cp, ncols, n, 2, op, n, cnt.vp,
                                      mgv.launch(0, 1024), st.vp))
            st.sync()
            times.append(t.durations_ms()[0])
        c = cnt.download(np.uint64, 2)
        ms = min(times[1:])
        algo = n * bytes_per_row
        gbps = algo / (ms / 1e3) / 1e9
        traffic = traffic_from_profiles(name)
        out[name] = {"ms": ms, "grows_per_s": n / (ms / 1e3) / 1e9, "gbps_algorithmic": gbps,
                     "frac": gbps / HBM_PEAK_GBPS, "bytes_per_row": bytes_per_row, "traffic": traffic,
                     "traffic_gbps": traffic / (ms / 1e3) / 1e9 if traffic else None,
                     "rows_per_parent": [int(v) for v in c], "kept_rows_ok": int(c.sum()) == want}
        log(f"F3 {name}: {ms:.4f} ms, {gbps:.0f} GB/s algorithmic ({gbps / HBM_PEAK_GBPS:.0%}), "
            f"rows {c.tolist()} ok {int(c.sum()) == want}")
    for b in (x, pred, cnt, res, *outs, *extra_outs):
        b.close()
    return out


# ------------------------------------------------------------------ C3 skew

def leg_c3(args, st):
    """Fused select -> join -> GROUP BY with skew: R has 2^28 rows {c uniform
    in [0, 2^31), k = Zipf(1.0) rank over 2^22 build keys mapped through a
    seeded permutation, v in [0, 2^16)}; S has 2^22 unique keys (a seeded
    permutation) with g in [0, 2^10).  GROUP BY S.g: COUNT, SUM(R.v) WHERE
    R.c <= t - 1 for selectivities 1/10/50/90 %.  The direct, the
    partitioned (L2) and the LDS-slice paths are timed; checks: the group
    arrays of all paths agree, their
    totals agree with a probe-sum kernel (COUNT, and SUM(R.v) = SUM(R.v +
    S.g) - sum_g g * COUNT_g), and the first 2^22 rows agree with the oracle."""
    nr, ns, G = 1 << 28, 1 << 22, 1 << 10
    rc = gen_bits(nr, mgv.seed(3, 0), 0, 31, st.vp)
    T = C.c_void_p()
    mgv.call("mgv_zipf_table_create", C.byref(T), ns, 1.0)
    rk = mgv.DeviceBuffer(4 * nr)
    mgv.call("mgv_gen_zipf", rk.vp, nr, mgv.seed(3, 1), 0, T, ns, mgv.seed(3, 2), 0, st.vp)
    rv = gen_bits(nr, mgv.seed(3, 3), 0, 16, st.vp)
    sk = gen_perm(ns, ns, mgv.seed(3, 4), 0, st.vp)
    sg = gen_bits(ns, mgv.seed(3, 5), 0, 10, st.vp)
    ht = mgv.HashTable(ns, 0.5)
    res = mgv.ResultBuffer(2)
    mgv.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, sk.vp, sg.vp, mgv.MGV_PAYLOAD_COLUMN, None, 0, ns, 0,
             res.vp, mgv.launch(0, 256), st.vp)
    st.sync()
    assert res.read(0).errors == 0
    gbuf = mgv.DeviceBuffer(16 * G)
    ws = mgv.partition_workspace(nr, 64, 2)
    need = C.c_uint64()
    mgv.call("mgv_probe_groupby_lds_workspace_size", nr, ht.nbuckets, 32, C.byref(need))
    wsl = mgv.DeviceBuffer(need.value)
    L = mgv.launch(0, 256, 1)
    Llds = mgv.Launch(0, 1024, -1, 2 << 8)
    out = {"probe_rows": nr, "build_rows": ns, "groups": G, "zipf_s": 1.0, "selectivity": {}}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    orc = oracle_py.Oracle()
    M = 1 << 22
    h = {k: b.download(np.int32, M if k.startswith("r") else ns) for k, b in
         (("rc", rc), ("rk", rk), ("rv", rv), ("sk", sk), ("sg", sg))}
    otab = orc.table(h["sk"], h["sg"])
    for sel, t in ((0.01, 21474836), (0.10, 214748365), (0.50, 1073741824), (0.90, 1932735283)):
        bound = t - 1
        paths = {}
        for name in ("direct", "part64", "lds32"):
            def run():
                mgv.call("mgv_fill_u64", gbuf.vp, 2 * G, 0, st.vp)
                if name == "direct":
                    mgv.call("mgv_hashjoin_probe_groupby", ht.buf.vp, ht.nbuckets, rc.vp, bound, rk.vp, rv.vp, nr,
                             gbuf.vp, gbuf.at(8 * G), G, L, st.vp)
                elif name == "lds32":
                    mgv.call("mgv_hashjoin_probe_groupby_lds", ht.buf.vp, ht.nbuckets, rc.vp, bound, rk.vp, rv.vp,
                             nr, gbuf.vp, gbuf.at(8 * G), G, 32, wsl.vp, wsl.nbytes, Llds, st.vp)
                else:
                    mgv.call("mgv_hashjoin_probe_groupby_partitioned", ht.buf.vp, ht.nbuckets, rc.vp, bound, rk.vp,
                             rv.vp, nr, gbuf.vp, gbuf.at(8 * G), G, 64, ws.vp, ws.nbytes, L, st.vp)
            times = []
            for rep in range(3):
                tm = Timer(st.vp)
                tm.around(run)
                st.sync()
                times.append(tm.durations_ms()[0])
            paths[name] = (min(times[1:]), gbuf.download(np.uint64, G), gbuf.download(np.int64, G, 8 * G))
        cnt, gs = paths["direct"][1], paths["direct"][2]
        same = all(bool(np.array_equal(cnt, paths[x][1]) and np.array_equal(gs, paths[x][2]))
                   for x in ("part64", "lds32"))
        # probe-sum totals: COUNT and SUM(R.v + S.g)
        res.clear(st.vp)
        mgv.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, rc.vp, bound, rk.vp, rv.vp, nr, 0, 0, res.slot(1), L,
                 st.vp)
        st.sync()
        ps = res.read(1)
        gsum_g = int((np.arange(G, dtype=np.int64) * cnt.astype(np.int64)).sum())
        totals_ok = int(cnt.sum()) == ps.count and int(gs.sum()) == ps.sum - gsum_g
        # oracle on the first M rows
        wc, wsum = orc.probe_groupby(otab, h["rk"], h["rv"], h["rc"], bound, G)
        mgv.call("mgv_fill_u64", gbuf.vp, 2 * G, 0, st.vp)
        mgv.call("mgv_hashjoin_probe_groupby", ht.buf.vp, ht.nbuckets, rc.vp, bound, rk.vp, rv.vp, M, gbuf.vp,
                 gbuf.at(8 * G), G, L, st.vp)
        st.sync()
        oracle_ok = bool(np.array_equal(gbuf.download(np.uint64, G), wc) and
                         np.array_equal(gbuf.download(np.int64, G, 8 * G), wsum))
        g8 = 1.0 - (1.0 - sel) ** 8
        algo = nr * (4 + 8 * g8 + 8 * sel)
        best = min(paths, key=lambda n: paths[n][0])
        ms = paths[best][0]
        out["selectivity"][f"{int(sel * 100)}%"] = {
            "bound": bound, "best_path": best, "ms": ms, "grows_per_s": nr / (ms / 1e3) / 1e9,
            "algorithmic_gbps": algo / (ms / 1e3) / 1e9, "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBPS,
            "ms_by_path": {n: paths[n][0] for n in paths}, "matches": int(cnt.sum()), "sum_v": int(gs.sum()),
            "paths_equal": same, "totals_equal_probe_sum": bool(totals_ok), "oracle_equal_first_2^22": oracle_ok}
        log(f"C3 s={sel:.0%}: {best} {ms:.3f} ms ({nr / ms / 1e6:.1f} Grows/s), paths equal {same}, "
            f"totals {totals_ok}, oracle {oracle_ok}")
        if sel == 0.50 and args.cpu:
            cpu = cpu_baseline_c3(args, st, rc, rk, rv, h["sk"], h["sg"], nr, bound, G, (cnt, gs))
            out["cpu_s50"] = cpu
            out["selectivity"]["50%"]["vs_cpu_all_threads"] = (nr / (ms / 1e3) / 1e9) / cpu["grows_per_s"]
    orc.free(otab)
    for b in (rc, rk, rv, sk, sg, ht.buf, res, gbuf, ws, wsl):
        b.close()
    mgv.load().mgv_free(T)
    return out


def cpu_baseline_c3(args, st, rc, rk, rv, sk, sg, nr, bound, G, gpu):
    """CPU baseline of C3 at one selectivity: the oracle's chained table over
    S (threaded build), then select -> probe -> GROUP BY S.g over all R rows
    on all threads, and over the first 2^24 rows on one thread; the group
    arrays of all rows must equal the GPU's."""
    orc = oracle()
    st.sync()
    h = {k: b.download(np.int32, nr) for k, b in (("rc", rc), ("rk", rk), ("rv", rv))}
    T = cpu_threads(args)
    t0 = time.perf_counter()
    tab = orc.table(sk, sg, threads=T)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    wc, wsum = orc.probe_groupby(tab, h["rk"], h["rv"], h["rc"], bound, G, threads=T)
    t_all = time.perf_counter() - t0
    M = min(nr, 1 << 24)
    t0 = time.perf_counter()
    orc.probe_groupby(tab, h["rk"][:M], h["rv"][:M], h["rc"][:M], bound, G)
    t_one = time.perf_counter() - t0
    orc.free(tab)
    equal = bool(np.array_equal(wc, gpu[0]) and np.array_equal(wsum, gpu[1]))
    log(f"C3 cpu (s=50%): {T} threads {t_all:.2f} s for {nr} rows, 1 thread {t_one:.2f} s for {M} rows; "
        f"equal to GPU {equal}")
    return {"threads": T, "build_s": t_build, "probe_groupby_s": t_all, "rows": nr, "grows_per_s": nr / t_all / 1e9,
            "one_thread": {"rows": M, "s": t_one, "grows_per_s": M / t_one / 1e9}, "equals_gpu": equal}


# ------------------------------------------------------------------ C4 batch

def tuned(workload: str):
    """Autotuned launch parameters (multigpu-vector_amd/tuned.json), if any."""
    p = os.path.join(ROOT, "multigpu-vector_amd", "tuned.json")
    try:
        with open(p) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def leg_c4(args, st):
    """256 select -> join -> SUM queries: one hipGraph (clear + one persistent
    launch for the builds + one for the probes), the same graph with the
    large probes routed through LDS table slices on side streams, and the
    plain path (one launch per vector per query).  All are timed by wall
    clock from issue to completion; each query's (COUNT, SUM) must agree
    between the paths, and with the oracle for the small queries."""
    import workloads
    t = tuned("c4") or {}
    V = int(t.get("vector_rows", 16384))
    blk = int(t.get("block", 256))
    compact = int(t.get("compact", 1))
    lf = float(t.get("load_factor", 0.5))
    shapes = workloads.c4_shapes(256)
    batch = workloads.JoinBatch(shapes, lf, st.vp)
    st.sync()
    log(f"C4: {len(shapes)} queries, {batch.r_rows} probe rows, {batch.s_rows} build rows, V={V} block={blk} "
        f"compact={compact} LF={lf}")
    L = mgv.launch(0, blk, compact)
    g = batch.capture_fused(V, L, st.vp)

    def run_graph():
        mgv.call("mgv_graph_launch", g, st.vp)

    def timed(fn, reps):
        walls, evs = [], []
        for _ in range(reps):
            a, b = mgv.Event(), mgv.Event()
            st.sync()
            t0 = time.perf_counter()
            a.record(st.vp)
            fn()
            b.record(st.vp)
            st.sync()
            walls.append(time.perf_counter() - t0)
            evs.append(a.elapsed_ms(b))
            a.close()
            b.close()
        return walls, evs

    timed(run_graph, 2)
    bw, be = timed(run_graph, args.c4_reps)
    fused = batch.results()
    assert int(fused["build_errors"].sum()) == 0
    mgv.call("mgv_graph_destroy", g)
    # routed: the large probes through LDS table slices, in the same graph
    thr = int(t.get("route_threshold", 1 << 20))
    nst = int(t.get("route_streams", 4))
    fan = int(t.get("route_fanout", 0))
    fan = fan if fan > 1 else 0   # 1 = the tuner kept the fused form: time the one-level LDS route
    g = batch.capture_routed(V, L, st.vp, thr, nst, fan)
    timed(run_graph, 2)
    rw, re_ = timed(run_graph, args.c4_reps)
    routed = batch.results()
    routed_same = bool(np.array_equal(routed["count"], fused["count"]) and
                       np.array_equal(routed["sum"], fused["sum"]))
    nbig = len(batch._routed(V, thr, nst, fan)["big"])
    mgv.call("mgv_graph_destroy", g)   # before any load-factor change frees what the graph uses
    log(f"C4 routed ({nbig} queries >= {thr} rows through LDS slices on {nst} streams): {min(rw) * 1e3:.2f} ms "
        f"wall, equal to fused: {routed_same}")
    # XCD-sliced: the large probes' rows scattered into table-part regions,
    # region r probed on XCD r mod 8 (mgv_batch_run_sliced), in one graph
    smin, sreg = 1 << 16, 8
    gs = batch.capture_sliced(V, L, st.vp, min_rows=smin, nregions=sreg)

    def run_sliced_graph():
        mgv.call("mgv_graph_launch", gs, st.vp)

    timed(run_sliced_graph, 2)
    sw_, se = timed(run_sliced_graph, args.c4_reps)
    sliced = batch.results()
    sliced_same = bool(np.array_equal(sliced["count"], fused["count"]) and
                       np.array_equal(sliced["sum"], fused["sum"]))
    nsl = batch._sliced(V, smin, sreg)["queries"]
    mgv.call("mgv_graph_destroy", gs)
    log(f"C4 sliced ({nsl} queries >= {smin} rows in {sreg} regions each): {min(sw_) * 1e3:.2f} ms wall, "
        f"equal to fused: {sliced_same}")
    # LDS batch: the large probes' rows scattered by table slice, each
    # (query, slice) probed from LDS (mgv_batch_run_lds), in one graph; the
    # slices hold half as many buckets at load factor 0.5, so it is timed at
    # the tuned load factor and at 0.5 (the tables are rebuilt in the graph)
    tl = t.get("lds_batch", {})
    lmin = int(tl.get("min_rows", 1 << 14))
    lV = int(tl.get("vector_rows", 1024))   # vector rows of the direct (small) queries
    ldsb = {}
    for lfl in sorted({lf, float(tl.get("load_factor", 0.5))}):
        if batch.lf != lfl:
            batch.set_load_factor(lfl)
        gl = batch.capture_ldsb(lV, L, st.vp, min_rows=lmin)

        def run_ldsb_graph():
            mgv.call("mgv_graph_launch", gl, st.vp)

        timed(run_ldsb_graph, 2)
        lw, le = timed(run_ldsb_graph, args.c4_reps)
        r = batch.results()
        same = bool(np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"]))
        mgv.call("mgv_graph_destroy", gl)
        # fresh_tables: the LDS-built tables lie in one span, left uncleared
        ldsb[lfl] = (lw, le, same, batch._ldsb(lV, lmin)["queries"], batch._ldsb(lV, lmin, True)["span"] is not None)
        log(f"C4 LDS batch at load factor {lfl} ({ldsb[lfl][3]} queries >= {lmin} rows, V={lV}): {min(lw) * 1e3:.2f} ms "
            f"wall, equal to fused: {same}")
    if batch.lf != lf:
        batch.set_load_factor(lf)
    ldsb_lf = min(ldsb, key=lambda k: min(ldsb[k][0]))
    lw, le, ldsb_same, nlq, lfresh = ldsb[ldsb_lf]
    ldsb_same = all(v[2] for v in ldsb.values())
    plain_out = {}
    plain_res = None
    for pv in ((4096, 65536) if args.c4_plain else ()):
        Lp = mgv.launch(0, 256, 1)
        timed(lambda: batch.issue_plain(pv, Lp, st.vp), 1)
        pw, pe = timed(lambda: batch.issue_plain(pv, Lp, st.vp), max(1, args.c4_reps // 2))
        r = batch.results()
        same = bool(np.array_equal(r["count"], fused["count"]) and np.array_equal(r["sum"], fused["sum"]))
        plain_out[str(pv)] = {"wall_ms": min(pw) * 1e3, "gpu_ms": min(pe), "queries_per_s": len(shapes) / min(pw),
                              "launches": 2 * (batch.items(pv)), "results_equal_fused": same}
        log(f"C4 plain V={pv}: {min(pw) * 1e3:.2f} ms wall, equal to fused: {same}")
        plain_res = r
    best_plain = max(plain_out.values(), key=lambda x: x["queries_per_s"]) if plain_out else None
    # oracle check of the smaller queries (host copies of their columns)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    orc = oracle_py.Oracle()
    checked, ok = 0, True
    for i, x in enumerate(shapes):
        if x.P > (1 << 17):
            continue
        h = batch.host_columns(i)
        tb = orc.table(h["sk"], h["sw"])
        want = orc.probe(tb, h["k"], h["v"], pred=h["c"], bound=x.bound)
        orc.free(tb)
        ok &= (int(fused["count"][i]), int(fused["sum"][i])) == (want["count"], want["sum"])
        checked += 1
    paths = {"fused": (bw, be), "routed": (rw, re_), "sliced": (sw_, se), "lds_batch": (lw, le)}
    qps = {k: len(shapes) / min(w) for k, (w, _) in paths.items()}
    fused_qps, routed_qps = qps["fused"], qps["routed"]
    best_path = max(qps, key=qps.get)
    best_w, best_e = paths[best_path]
    algo = batch.algorithmic_bytes(ldsb_lf if best_path == "lds_batch" else None)
    out = {
        "queries": len(shapes), "probe_rows": batch.r_rows, "build_rows": batch.s_rows,
        "tuned": {"vector_rows": V, "block": blk, "compact": compact, "load_factor": lf, "from_file": bool(t)},
        "load_factor_of_best_path": ldsb_lf if best_path == "lds_batch" else lf,
        "batched": {"wall_ms": min(best_w) * 1e3, "gpu_ms": min(best_e), "queries_per_s": qps[best_path],
                    "path": best_path,
                    "algorithmic_gbps": algo / (min(best_e) / 1e3) / 1e9,
                    "frac": algo / (min(best_e) / 1e3) / 1e9 / HBM_PEAK_GBPS, "work_items": batch.items(V)},
        "fused_only": {"wall_ms": min(bw) * 1e3, "gpu_ms": min(be), "queries_per_s": fused_qps},
        "routed": {"wall_ms": min(rw) * 1e3, "gpu_ms": min(re_), "queries_per_s": routed_qps,
                   "lds_queries": nbig, "threshold_rows": thr, "side_streams": nst, "fanout": fan or "one level"},
        "sliced": {"wall_ms": min(sw_) * 1e3, "gpu_ms": min(se), "queries_per_s": qps["sliced"],
                   "sliced_queries": nsl, "min_rows": smin, "regions": sreg},
        "lds_batch": {"wall_ms": min(lw) * 1e3, "gpu_ms": min(le), "queries_per_s": qps["lds_batch"],
                      "lds_queries": nlq, "min_rows": lmin, "vector_rows": lV, "load_factor": ldsb_lf,
                      "fresh_tables": lfresh,
                      "wall_ms_by_load_factor": {str(k): min(v[0]) * 1e3 for k, v in ldsb.items()}},
        "plain": plain_out,
        "speedup_vs_best_plain": qps[best_path] / best_plain["queries_per_s"] if best_plain else None,
        "parity": {"fused_equals_plain": (all(v["results_equal_fused"] for v in plain_out.values())
                                          if plain_out else None),
                   "routed_equals_fused": routed_same, "sliced_equals_fused": sliced_same,
                   "lds_batch_equals_fused": ldsb_same,
                   "oracle_checked_queries": checked, "oracle_equal": bool(ok)},
        "total_count": int(fused["count"].sum()), "total_sum": int(fused["sum"].sum()),
    }
    sp = out["speedup_vs_best_plain"]
    log(f"C4 batched ({best_path}) {min(best_w) * 1e3:.2f} ms wall ({qps[best_path]:.0f} q/s), "
        f"{'x%.2f vs best plain' % sp if sp else 'plain path not run'}; oracle {checked} queries ok={ok}")
    del plain_res
    batch.close()
    return out


# ------------------------------------------------------------------ C5 serving

def leg_c5(args, st):
    """Serving: requests over 64 GiB of resident int32 columns (70 %
    select -> join -> SUM over build tables of 2^16..2^26 rows, 30 % GROUP BY
    over 2^16 Zipf(0.8) groups), admitted against a 16 GiB HBM budget.
    Three runs of the same requests: saturated (all arrive at t=0) to find
    the capacity, Poisson arrivals at 70 % of that capacity (the reported
    latencies), Poisson arrivals at 95 % of it (batches merge), and one
    request at a time (the alone reference): every request's (COUNT, SUM,
    group digest) must be identical in all four.  Join probe keys follow
    Zipf(1.0) over their build keys (the skewed select -> join shape)."""
    import workloads
    t0 = time.perf_counter()
    tables = workloads.ServeTables(build_logs=range(16, 27), probe_cap=1 << 29,
                                   group_rows=(1 << 30,) * 5 + (1 << 29,), ngroups=1 << 16, zipf_s=0.8,
                                   join_zipf_s=1.0, stream=st.vp)
    st.sync()
    log(f"C5: {tables.bytes / 2**30:.1f} GiB of resident columns generated in {time.perf_counter() - t0:.1f} s")
    n = args.c5_requests
    budget = 16 << 30
    V = 65536
    reqs0, desc = tables.requests(n, 0.0, seed=0xC5, arrivals=False)
    sat, ss = workloads.serve(reqs0, n, budget, max_inflight=4, max_batch=64, vector_rows=V)
    log(f"C5 saturated: {ss.queries_per_s:.1f} q/s, {ss.batches} batches")
    alone, sa = workloads.serve(reqs0, n, budget, max_inflight=1, max_batch=1, vector_rows=V)
    log(f"C5 one at a time: {sa.queries_per_s:.1f} q/s")
    rate = 0.7 * ss.queries_per_s
    reqs, _ = tables.requests(n, rate, seed=0xC5)
    mix, sm = workloads.serve(reqs, n, budget, max_inflight=4, max_batch=64, vector_rows=V)
    log(f"C5 Poisson {rate:.1f} q/s: p50 {sm.p50_s * 1e3:.2f} ms p99 {sm.p99_s * 1e3:.2f} ms, "
        f"{sm.queries_per_s:.1f} q/s, mean batch {sm.mean_batch:.1f}")
    # near capacity: requests keep arriving while others run, so batches merge
    rate_hi = 0.95 * ss.queries_per_s
    reqs_hi, _ = tables.requests(n, rate_hi, seed=0xC5)
    mix_hi, sh = workloads.serve(reqs_hi, n, budget, max_inflight=4, max_batch=64, vector_rows=V)
    log(f"C5 Poisson {rate_hi:.1f} q/s: p50 {sh.p50_s * 1e3:.2f} ms p99 {sh.p99_s * 1e3:.2f} ms, "
        f"{sh.queries_per_s:.1f} q/s, mean batch {sh.mean_batch:.1f}")

    def trip(r):
        return [(r[i].count, r[i].sum, r[i].digest) for i in range(n)]

    same = trip(mix) == trip(alone) == trip(sat) == trip(mix_hi)
    errs = sum(mix[i].errors for i in range(n))
    kinds = [d["kind"] for d in desc]
    lat_join = sorted(mix[i].latency_s for i in range(n) if kinds[i] == "join")
    lat_grp = sorted(mix[i].latency_s for i in range(n) if kinds[i] == "groupby")

    def pct(v, p):
        return v[min(len(v) - 1, max(0, int(np.ceil(p * len(v))) - 1))] * 1e3 if v else None

    out = {
        "requests": n, "join_requests": kinds.count("join"), "groupby_requests": kinds.count("groupby"),
        "resident_column_gib": tables.bytes / 2**30, "hbm_budget_gib": budget / 2**30, "vector_rows": V,
        "max_inflight": 4, "max_batch": 64,
        "poisson": {"rate_qps": rate, "queries_per_s": sm.queries_per_s, "p50_ms": sm.p50_s * 1e3,
                    "p99_ms": sm.p99_s * 1e3, "max_ms": sm.max_s * 1e3, "mean_ms": sm.mean_s * 1e3,
                    "batches": sm.batches, "mean_batch": sm.mean_batch, "max_inflight_seen": sm.max_inflight_seen,
                    "peak_budget_gib": sm.peak_bytes / 2**30,
                    "join_p50_ms": pct(lat_join, 0.5), "join_p99_ms": pct(lat_join, 0.99),
                    "groupby_p50_ms": pct(lat_grp, 0.5), "groupby_p99_ms": pct(lat_grp, 0.99)},
        "poisson_near_capacity": {"rate_qps": rate_hi, "queries_per_s": sh.queries_per_s, "p50_ms": sh.p50_s * 1e3,
                                  "p99_ms": sh.p99_s * 1e3, "mean_batch": sh.mean_batch, "batches": sh.batches,
                                  "max_inflight_seen": sh.max_inflight_seen},
        "saturated": {"queries_per_s": ss.queries_per_s, "p50_ms": ss.p50_s * 1e3, "p99_ms": ss.p99_s * 1e3,
                      "mean_batch": ss.mean_batch},
        "results_through_completion_ring": int(sm.ring_results + sh.ring_results + ss.ring_results
                                               + sa.ring_results),
        "one_at_a_time": {"queries_per_s": sa.queries_per_s, "p50_ms": sa.p50_s * 1e3, "p99_ms": sa.p99_s * 1e3},
        "parity": {"mix_equals_alone_equals_saturated": bool(same), "build_or_group_errors": int(errs)},
    }
    tables.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--block", type=int, default=0, help="probe workgroup size (0 = tuned, else 256)")
    ap.add_argument("--coop", type=int, default=-1, help="4-lane cooperative probe (-1 = tuned)")
    ap.add_argument("--load-factor", type=float, default=0.5)
    ap.add_argument("--probe", default="auto", choices=["auto", "direct", "part", "lds", "ldsc", "ldsk"],
                    help="probe path (auto = time the candidates, keep the fastest)")
    ap.add_argument("--fanout", type=int, default=-1, help="fan-out of the partitioned paths (with --probe)")
    ap.add_argument("--scatter", type=int, default=-1, help="scatter workgroup shape MGV_SCATTER_* (-1 = tuned)")
    ap.add_argument("--legs", default="c1,c3,c4,c5,copy,f3,f1")
    ap.add_argument("--c5-requests", type=int, default=300)
    ap.add_argument("--c4-reps", type=int, default=5)
    ap.add_argument("--c4-plain", type=int, default=1, help="time the plain per-vector path in the C4 leg (0 = skip)")
    ap.add_argument("--cpu", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=1 << 25, help="R rows of the one-thread C2 figure")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    mgv.load()
    mgv.init(0)
    p = mgv.props()
    log(f"device {p.name.decode()} {p.arch.decode()} CUs={p.cus} HBM={p.hbm_bytes >> 30} GiB")
    st = mgv.Stream()
    c2 = leg_c2(args, st)
    extra = {}
    legs = [x for x in args.legs.split(",") if x]
    if "c1" in legs:
        extra["c1"] = leg_c1(args, st)
    if "c3" in legs:
        extra["c3"] = leg_c3(args, st)
    if "c4" in legs:
        mgv.load_ops()
        extra["c4"] = leg_c4(args, st)
    if "copy" in legs:
        extra["copy"] = leg_copy(args, st)
    if "f3" in legs:
        extra["f3"] = leg_f3(args, st)
    if "f1" in legs:
        mgv.load_ops()
        extra["f1"] = leg_f1(args, st)
    if "c5" in legs:
        mgv.load_ops()
        extra["c5"] = leg_c5(args, st)

    cpu = None
    if args.cpu and "cpu_c2" in c2:
        c = c2["cpu_c2"]
        cpu = {"value": c["grows_per_s"], "unit": "Grows/s", "cores": c["threads"], "kind": "port",
               "sample": f"oracle chained hash join (the reference's CPU algorithm, main8.cu:84-103): threaded "
                         f"build of all {1 << 24} S rows, probe of all {1 << 28} R rows on {c['threads']} threads; "
                         f"one-thread figure on the first {c['one_thread']['probe_rows']} R rows",
               **cpu_info(), "c2": c}
        if "c1" in extra and "cpu" in extra["c1"]:
            cpu["c1"] = extra["c1"]["cpu"]
        if "c3" in extra and "cpu_s50" in extra["c3"]:
            cpu["c3_s50"] = extra["c3"]["cpu_s50"]

    line = {
        "metric": "probe Grows/s & %HBM peak; batched vs plain queries/s at bit-exact, p99 lat",
        "value": c2["value"],
        "unit": "Grows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c2["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Philox4x32-10 columns generated on the GPU)",
        "config": {"workload": "C2 hash join + aggregate: probe 2^28 R rows against a resident open-addressing "
                               "table of 2^24 S rows at load factor 0.5; COUNT(*), SUM(R.v+S.w) int64",
                   "build_rows": 1 << 24, "probe_rows": 1 << 28, "load_factor": args.load_factor,
                   "probe_path": c2["roofline"]["probe_path"], "fanout": c2["roofline"]["fanout"],
                   "parallelism": "1 GPU"},
        "roofline": c2["roofline"],
        "cpu_baseline": cpu,
        "probe_kernel_ms": {k: c2[k] for k in ("kernel_ms_avg", "kernel_ms_min", "kernel_ms_median")},
        "probe_path_sweep_ms": c2["probe_path_sweep_ms"],
        "build": c2["build"],
        "result": c2["result"],
        "device": {"name": p.name.decode(), "arch": p.arch.decode(), "cus": p.cus},
    }
    line.update(extra)
    failed = parity_failures(line)
    line["parity_ok"] = not failed
    print(json.dumps(line), flush=True)
    if failed:
        log("PARITY FAILURE: " + ", ".join(failed))
        sys.exit(3)


def parity_failures(line):
    """Names of the parity flags of the JSON line that are false."""
    flags = {
        "c2 timed path = closed form": line["result"]["timed_path_equals_closed_form"],
        "c2 every candidate = closed form": line["result"]["every_candidate_equals_closed_form"],
    }
    cpu = line.get("cpu_baseline") or {}
    if "c2" in cpu:
        flags["c2 cpu = gpu"] = cpu["c2"]["equals_gpu"]
        flags["c2 cpu one-thread sample = gpu"] = cpu["c2"]["one_thread_sample_equals_gpu"]
    for k in ("one_thread", "all_threads"):
        if k in cpu.get("c1", {}):
            flags[f"c1 cpu {k} = gpu"] = cpu["c1"][k]["equals_gpu"]
    if "c3_s50" in cpu:
        flags["c3 cpu = gpu"] = cpu["c3_s50"]["equals_gpu"]
    c1 = line.get("c1")
    if c1 and "c1_fused_batch" in c1:
        flags["c1 fused batch = per-query calls"] = c1["c1_fused_batch"]["results_equal_per_query_calls"]
    for sel, v in (line.get("c3") or {}).get("selectivity", {}).items():
        flags[f"c3 {sel} paths equal"] = v["paths_equal"]
        flags[f"c3 {sel} totals"] = v["totals_equal_probe_sum"]
        flags[f"c3 {sel} oracle"] = v["oracle_equal_first_2^22"]
    c4 = line.get("c4")
    if c4:
        if c4["parity"]["fused_equals_plain"] is not None:
            flags["c4 fused = plain"] = c4["parity"]["fused_equals_plain"]
        flags["c4 routed = fused"] = c4["parity"]["routed_equals_fused"]
        flags["c4 sliced = fused"] = c4["parity"]["sliced_equals_fused"]
        flags["c4 LDS batch = fused"] = c4["parity"]["lds_batch_equals_fused"]
        flags["c4 oracle"] = c4["parity"]["oracle_equal"]
    for k, v in (line.get("f3") or {}).items():
        flags[f"f3 {k} kept rows"] = v["kept_rows_ok"]
    if line.get("f1"):
        flags["f1 host columns = device columns"] = line["f1"]["sums_equal"]
    c5 = line.get("c5")
    if c5:
        flags["c5 mix = alone = saturated"] = c5["parity"]["mix_equals_alone_equals_saturated"]
        flags["c5 no build/group errors"] = c5["parity"]["build_or_group_errors"] == 0
    return [k for k, v in flags.items() if not v]


if __name__ == "__main__":
    main()
