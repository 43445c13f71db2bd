"""Seeded query batches over device-resident columns, shared by bench.py, the
autotuner and the serving loop.

A batch is a list of select -> hash-join -> SUM queries (the reference's
select / hashjoin_builder / hashjoin / aggregation chain, src/main9.cu:260-367,
with int64 sums): query q has a probe table R_q of P_q rows {c, k, v} and a
build table S_q of B_q = P_q / 16 unique keys {k, w}.  The query computes
COUNT(*), SUM(R.v + S.w) over R JOIN S ON k WHERE R.c <= bound_q.  All columns
come from the device Philox generator, so a query's inputs depend only on
(seed, query index) and are identical whether it runs alone or in a batch.

Columns of all queries live in a few large device arenas (one per column),
with every query's slice 256-byte aligned; hash tables share one arena too,
so one fill kernel clears every table of the batch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

import mgv

ALIGN_ROWS = 64          # 256-byte aligned column slices
INT32_MAX = 2**31 - 1


def _align(x: int, a: int) -> int:
    return (x + a - 1) // a * a


@dataclass
class JoinQuery:
    P: int                    # probe rows
    B: int                    # build rows (unique keys 0..B-1)
    sel: float                # filter selectivity
    bound: int                # c <= bound
    stream: int               # Philox stream of the query's columns
    r_off: int = 0            # row offset of R_q in the probe arenas
    s_off: int = 0            # row offset of S_q in the build arenas
    t_off: int = 0            # byte offset of the table in the table arena
    nbuckets: int = 0
    extra: dict = field(default_factory=dict)


def c4_shapes(nq: int = 256, seed: int = 0xC4, lo: float = 12.0, hi: float = 24.0) -> List[JoinQuery]:
    """Query shapes: P = floor(2^u), u ~ U[lo, hi]; B = P/16; selectivity
    ~ U[0.05, 0.95]; bound = floor(sel * 2^31) - 1 (c is uniform in [0, 2^31))."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(lo, hi, nq)
    sel = rng.uniform(0.05, 0.95, nq)
    out = []
    for q in range(nq):
        P = int(np.floor(2.0 ** u[q]))
        B = max(16, P // 16)
        bound = min(INT32_MAX, int(np.floor(sel[q] * 2**31)) - 1)
        out.append(JoinQuery(P=P, B=B, sel=float(sel[q]), bound=bound, stream=q))
    return out


class JoinBatch:
    """Device columns, hash tables and query descriptors of a join batch."""

    CFG = 4   # seed config of the batch columns

    def __init__(self, queries: List[JoinQuery], load_factor: float = 0.5, stream=None):
        self.q = queries
        self.lf = load_factor
        self.st = stream
        r_rows = s_rows = 0
        for x in queries:
            x.r_off, r_rows = r_rows, r_rows + _align(x.P, ALIGN_ROWS)
            x.s_off, s_rows = s_rows, s_rows + _align(x.B, ALIGN_ROWS)
        self.r_rows, self.s_rows = r_rows, s_rows
        self.rc = mgv.DeviceBuffer(4 * max(1, r_rows))
        self.rk = mgv.DeviceBuffer(4 * max(1, r_rows))
        self.rv = mgv.DeviceBuffer(4 * max(1, r_rows))
        self.sk = mgv.DeviceBuffer(4 * max(1, s_rows))
        self.sw = mgv.DeviceBuffer(4 * max(1, s_rows))
        self.tables = None
        self.nq = len(queries)
        self.res = mgv.ResultBuffer(2 * max(1, self.nq))   # [0, nq) builds, [nq, 2nq) probes
        sd = mgv.seed
        for x in queries:
            r4, s4 = 4 * x.r_off, 4 * x.s_off
            mgv.call("mgv_gen_bits", self.rc.at(r4), x.P, sd(self.CFG, 0), x.stream, 31, stream)
            mgv.call("mgv_gen_range", self.rk.at(r4), x.P, sd(self.CFG, 1), x.stream, 2 * x.B, stream)
            mgv.call("mgv_gen_bits", self.rv.at(r4), x.P, sd(self.CFG, 2), x.stream, 16, stream)
            mgv.call("mgv_gen_perm", self.sk.at(s4), x.B, x.B, sd(self.CFG, 3), x.stream, stream)
            mgv.call("mgv_gen_bits", self.sw.at(s4), x.B, sd(self.CFG, 4), x.stream, 16, stream)
        self.counter = mgv.DeviceBuffer(256)
        self._plans = {}
        self.set_load_factor(load_factor)

    def set_load_factor(self, load_factor: float):
        """(Re)allocates the hash tables for a load factor and rewrites the
        query descriptors; the columns are kept.  Tables lie in order of
        decreasing build size, so the tables an LDS batch builds (the builds
        between a threshold and the slice limit) are one contiguous span."""
        t_bytes = 0
        for x in sorted(self.q, key=lambda y: -y.B):
            nb, nbytes = mgv.ht_plan(x.B, load_factor)
            x.nbuckets = nb
            x.t_off, t_bytes = t_bytes, t_bytes + _align(nbytes, 256)
        self.lf, self.t_bytes = load_factor, t_bytes
        if self.tables is not None:
            self.tables.close()
        self.tables = mgv.DeviceBuffer(max(8, t_bytes))
        self._drop_plans()
        self.build_q = (mgv.Query * max(1, self.nq))()
        self.probe_q = (mgv.Query * max(1, self.nq))()
        for i, x in enumerate(self.q):
            slots = self.tables.ptr + x.t_off
            self.build_q[i] = mgv.Query(mgv.MGV_Q_BUILD, 0, None, self.sk.ptr + 4 * x.s_off,
                                        self.sw.ptr + 4 * x.s_off, x.B, slots, x.nbuckets, 0, None, None,
                                        self.res.ptr + 64 * i, mgv.MGV_PAYLOAD_COLUMN, 0, 0)
            self.probe_q[i] = mgv.Query(mgv.MGV_Q_PROBE_SUM, x.bound, self.rc.ptr + 4 * x.r_off,
                                        self.rk.ptr + 4 * x.r_off, self.rv.ptr + 4 * x.r_off, x.P, slots,
                                        x.nbuckets, 0, None, None, self.res.ptr + 64 * (self.nq + i), 0, 0, 0)

    def _drop_plans(self):
        for plan in self._plans.values():
            if isinstance(plan, dict) and ("sliced" in plan or "ldsb" in plan):
                for b in (plan["dq"], plan["di"], plan["dr"], plan["ws"]):
                    b.close()
                continue
            if isinstance(plan, dict):   # routed plan
                dq, di, _ = plan["small"]
                for b in [dq, di] + plan["ws"] + plan["streams"] + plan["events"]:
                    b.close()
                continue
            for dq, di, _ in plan:
                dq.close()
                di.close()
        self._plans = {}

    # -------------------------------------------------------------- common prologue
    def clear(self, stream=None):
        """Empty every table (one fill kernel) and zero every result."""
        s = stream if stream is not None else self.st
        mgv.call("mgv_fill_u64", self.tables.vp, self.t_bytes // 8, 0xFFFFFFFFFFFFFFFF, s)
        self.res.clear(s)

    # -------------------------------------------------------------- fused path
    def _plan(self, vector_rows: int):
        if vector_rows in self._plans:
            return self._plans[vector_rows]
        out = []
        for qarr in (self.build_q, self.probe_q):
            n = C.c_uint64()
            mgv.call("mgv_batch_plan", qarr, self.nq, vector_rows, None, 0, C.byref(n))
            items = (mgv.WorkItem * max(1, n.value))()
            mgv.call("mgv_batch_plan", qarr, self.nq, vector_rows, items, n.value, C.byref(n))
            dq = mgv.DeviceBuffer(C.sizeof(qarr)).upload(np.frombuffer(bytes(qarr), dtype=np.uint8))
            di = mgv.DeviceBuffer(max(16, C.sizeof(items))).upload(np.frombuffer(bytes(items), dtype=np.uint8))
            out.append((dq, di, n.value))
        self._plans[vector_rows] = out
        return out

    def issue_fused(self, vector_rows: int, launch: mgv.Launch, stream=None):
        """clear + one persistent launch for all builds + one for all probes."""
        s = stream if stream is not None else self.st
        (bq, bi, bn), (pq, pi, pn) = self._plan(vector_rows)
        self.clear(s)
        mgv.call("mgv_batch_run", bq.vp, bi.vp, bn, self.counter.vp, launch, s)
        mgv.call("mgv_batch_run", pq.vp, pi.vp, pn, self.counter.at(128), launch, s)

    def capture_fused(self, vector_rows: int, launch: mgv.Launch, stream) -> C.c_void_p:
        self._plan(vector_rows)
        mgv.call("mgv_graph_begin_capture", stream)
        self.issue_fused(vector_rows, launch, stream)
        g = C.c_void_p()
        mgv.call("mgv_graph_end_capture", stream, C.byref(g))
        return g

    # -------------------------------------------------------------- routed path
    def lds_fanout(self, x: JoinQuery, want: int = 0) -> int:
        """Level-1 fan-out of a query's LDS probe: `want` when the call takes
        it for this table (one level, or two when the table needs more
        slices), else the fan-out that cuts the table into slices of <= 2048
        buckets in one level (0 when its bucket count does not split)."""
        if want:
            b = C.c_uint64()
            if mgv.load().mgv_probe_lds_workspace_size(C.c_uint64(x.P), C.c_uint32(x.nbuckets), C.c_uint32(want),
                                                       C.byref(b)) == 0:
                return want
        f = 2
        while x.nbuckets // f > 2048:
            f *= 2
        return f if x.nbuckets % f == 0 and f <= 2048 else 0

    def _routed(self, vector_rows: int, threshold: int, nstreams: int, fanout: int = 0):
        """Plan of the routed batch: the fused items of the small probes
        (P < threshold) and the large probes' LDS calls; cached."""
        key = ("routed", vector_rows, threshold, nstreams, fanout)
        if key in self._plans:
            return self._plans[key]
        big = [i for i, x in enumerate(self.q) if x.P >= threshold and self.lds_fanout(x, fanout)]
        small = [i for i in range(self.nq) if i not in set(big)]
        sq = (mgv.Query * max(1, len(small)))()
        for j, i in enumerate(small):
            sq[j] = self.probe_q[i]
        n = C.c_uint64()
        mgv.call("mgv_batch_plan", sq, len(small), vector_rows, None, 0, C.byref(n))
        items = (mgv.WorkItem * max(1, n.value))()
        if small:
            mgv.call("mgv_batch_plan", sq, len(small), vector_rows, items, n.value, C.byref(n))
        dq = mgv.DeviceBuffer(C.sizeof(sq)).upload(np.frombuffer(bytes(sq), dtype=np.uint8))
        di = mgv.DeviceBuffer(max(16, C.sizeof(items))).upload(np.frombuffer(bytes(items), dtype=np.uint8))
        need = 0
        fan = {i: self.lds_fanout(self.q[i], fanout) for i in big}
        for i in big:
            x = self.q[i]
            b = C.c_uint64()
            mgv.call("mgv_probe_lds_workspace_size", x.P, x.nbuckets, fan[i], C.byref(b))
            need = max(need, b.value)
        ws = [mgv.DeviceBuffer(max(256, need)) for _ in range(nstreams)] if big else []
        streams = [mgv.Stream() for _ in range(nstreams)] if big else []
        events = [mgv.Event() for _ in range(nstreams + 1)] if big else []
        plan = {"small": (dq, di, n.value if small else 0), "big": big, "fan": fan, "ws": ws, "streams": streams,
                "events": events}
        self._plans[key] = plan
        return plan

    def issue_routed(self, vector_rows: int, launch: mgv.Launch, stream, threshold: int = 1 << 20,
                     nstreams: int = 4, fanout: int = 0):
        """clear + the fused build launch; then the small probes in one fused
        launch on `stream` while the large ones (P >= threshold, tables that
        split into LDS slices) run through mgv_hashjoin_probe_lds (level-1
        fan-out `fanout` where the table takes it, see lds_fanout) on
        nstreams side streams (forked and joined with events, so the whole
        batch can be captured into one graph).  Results land in the same
        slots as the fused path's."""
        (bq, bi, bn), _ = self._plan(vector_rows)
        plan = self._routed(vector_rows, threshold, nstreams, fanout)
        self.clear(stream)
        mgv.call("mgv_batch_run", bq.vp, bi.vp, bn, self.counter.vp, launch, stream)
        ev = plan["events"]
        if plan["big"]:
            ev[0].record(stream)
            for sd in plan["streams"]:
                mgv.call("mgv_stream_wait_event", sd.vp, C.c_void_p(ev[0].ptr))
            lds_launch = mgv.Launch(0, 1024, -1, 2 << 8)
            for j, i in enumerate(plan["big"]):
                x = self.q[i]
                k = j % len(plan["streams"])
                ws, sd = plan["ws"][k], plan["streams"][k]
                r4 = 4 * x.r_off
                mgv.call("mgv_hashjoin_probe_lds", self.tables.ptr + x.t_off, x.nbuckets, self.rc.at(r4), x.bound,
                         self.rk.at(r4), self.rv.at(r4), x.P, plan["fan"][i], ws.vp, ws.nbytes,
                         self.res.slot(self.nq + i), lds_launch, sd.vp)
        dq, di, pn = plan["small"]
        if pn:
            mgv.call("mgv_batch_run", dq.vp, di.vp, pn, self.counter.at(128), launch, stream)
        if plan["big"]:
            for k, sd in enumerate(plan["streams"]):
                ev[k + 1].record(sd.vp)
                mgv.call("mgv_stream_wait_event", stream, C.c_void_p(ev[k + 1].ptr))

    def capture_routed(self, vector_rows: int, launch: mgv.Launch, stream, threshold: int = 1 << 20,
                       nstreams: int = 4, fanout: int = 0) -> C.c_void_p:
        self._plan(vector_rows)
        self._routed(vector_rows, threshold, nstreams, fanout)
        mgv.call("mgv_graph_begin_capture", stream)
        self.issue_routed(vector_rows, launch, stream, threshold, nstreams, fanout)
        g = C.c_void_p()
        mgv.call("mgv_graph_end_capture", stream, C.byref(g))
        return g

    # -------------------------------------------------------------- XCD-sliced path
    def _sliced(self, vector_rows: int, min_rows: int, nregions: int = 0):
        """Plan of the sliced probe launch (mgv_batch_plan_sliced over the
        probe queries): device queries, items, regions and workspace; cached."""
        key = ("sliced", vector_rows, min_rows, nregions)
        if key in self._plans:
            return self._plans[key]
        plan = mgv.SlicedPlan()
        regions = (C.c_uint64 * (4 * max(1, self.nq)))()
        mgv.call("mgv_batch_plan_sliced", self.probe_q, self.nq, vector_rows, min_rows, nregions, None, 0, regions,
                 C.byref(plan))
        items = (mgv.WorkItem * max(1, plan.nitems))()
        mgv.call("mgv_batch_plan_sliced", self.probe_q, self.nq, vector_rows, min_rows, nregions, items, plan.nitems,
                 regions, C.byref(plan))
        dq = mgv.DeviceBuffer(C.sizeof(self.probe_q)).upload(np.frombuffer(bytes(self.probe_q), dtype=np.uint8))
        di = mgv.DeviceBuffer(max(16, C.sizeof(items))).upload(np.frombuffer(bytes(items), dtype=np.uint8))
        dr = mgv.DeviceBuffer(C.sizeof(regions)).upload(np.frombuffer(bytes(regions), dtype=np.uint8))
        ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
        out = {"sliced": True, "plan": plan, "dq": dq, "di": di, "dr": dr, "ws": ws,
               "queries": sum(1 for i in range(self.nq) if regions[4 * i + 2])}
        self._plans[key] = out
        return out

    def issue_sliced(self, vector_rows: int, launch: mgv.Launch, stream=None, min_rows: int = 1 << 20,
                     nregions: int = 0):
        """clear + the fused build launch + the XCD-sliced probe
        (mgv_batch_run_sliced: probes of at least min_rows rows scattered
        into nregions regions by table part, regions r = x mod 8 probed on
        XCD x)."""
        s = stream if stream is not None else self.st
        (bq, bi, bn), _ = self._plan(vector_rows)
        sp = self._sliced(vector_rows, min_rows, nregions)
        self.clear(s)
        mgv.call("mgv_batch_run", bq.vp, bi.vp, bn, self.counter.vp, launch, s)
        mgv.call("mgv_batch_run_sliced", sp["dq"].vp, sp["di"].vp, sp["dr"].vp, C.byref(sp["plan"]), sp["ws"].vp,
                 sp["ws"].nbytes, launch, s)

    def capture_sliced(self, vector_rows: int, launch: mgv.Launch, stream, min_rows: int = 1 << 20,
                       nregions: int = 0) -> C.c_void_p:
        self._plan(vector_rows)
        self._sliced(vector_rows, min_rows, nregions)
        mgv.call("mgv_graph_begin_capture", stream)
        self.issue_sliced(vector_rows, launch, stream, min_rows, nregions)
        g = C.c_void_p()
        mgv.call("mgv_graph_end_capture", stream, C.byref(g))
        return g

    # -------------------------------------------------------------- LDS batch path
    def _ldsb(self, vector_rows: int, min_rows: int, builds: bool = False):
        """Plan of the LDS batch (mgv_batch_plan_lds over the probe queries,
        or over the build queries): device queries, items, regions and
        workspace; cached."""
        key = ("ldsb", vector_rows, min_rows, builds)
        if key in self._plans:
            return self._plans[key]
        qarr = self.build_q if builds else self.probe_q
        plan = mgv.LdsBatchPlan()
        regions = (C.c_uint64 * (4 * max(1, self.nq)))()
        mgv.call("mgv_batch_plan_lds", qarr, self.nq, vector_rows, min_rows, None, 0, regions, C.byref(plan))
        items = (mgv.WorkItem * max(1, plan.nitems))()
        mgv.call("mgv_batch_plan_lds", qarr, self.nq, vector_rows, min_rows, items, plan.nitems, regions,
                 C.byref(plan))
        dq = mgv.DeviceBuffer(C.sizeof(qarr)).upload(np.frombuffer(bytes(qarr), dtype=np.uint8))
        di = mgv.DeviceBuffer(max(16, C.sizeof(items))).upload(np.frombuffer(bytes(items), dtype=np.uint8))
        dr = mgv.DeviceBuffer(C.sizeof(regions)).upload(np.frombuffer(bytes(regions), dtype=np.uint8))
        ws = mgv.DeviceBuffer(max(256, plan.workspace_bytes))
        out = {"ldsb": True, "plan": plan, "dq": dq, "di": di, "dr": dr, "ws": ws,
               "queries": sum(1 for i in range(self.nq) if regions[4 * i + 2]),
               "span": self._lds_span([i for i in range(self.nq) if regions[4 * i + 2]]) if builds else None}
        self._plans[key] = out
        return out

    def _lds_span(self, built):
        """Byte span [lo, hi) of the tables of the queries in built when they
        lie contiguously (no other table between them), else None."""
        if not built:
            return None
        offs = sorted((self.q[i].t_off for i in built))
        lo = offs[0]
        hi = max(self.q[i].t_off for i in built)
        hi = min((x.t_off for x in self.q if x.t_off > hi), default=self.t_bytes)
        inside = [x for x in self.q if lo <= x.t_off < hi]
        return (lo, hi) if len(inside) == len(built) else None

    def issue_ldsb(self, vector_rows: int, launch: mgv.Launch, stream=None, min_rows: int = 1 << 16,
                   lds_builds: bool = True, fresh: bool = True):
        """clear + the builds (through the LDS batch when lds_builds, else
        the fused build launch) + the LDS batch probe (mgv_batch_run_lds:
        probes of at least min_rows rows scattered by table slice, each
        (query, slice) probed from LDS).  With fresh, the LDS-built tables
        (one contiguous span) are not cleared: MGV_LAUNCH_FRESH_TABLES writes
        them whole."""
        s = stream if stream is not None else self.st
        (bq, bi, bn), _ = self._plan(vector_rows)
        lp = self._ldsb(vector_rows, min_rows)
        lb = self._ldsb(vector_rows, min_rows, builds=True) if lds_builds else None
        span = lb["span"] if (lb is not None and fresh) else None
        if span is None:
            self.clear(s)
        else:
            for a, b in ((0, span[0]), (span[1], self.t_bytes)):
                if b > a:
                    mgv.call("mgv_fill_u64", self.tables.ptr + a, (b - a) // 8, 0xFFFFFFFFFFFFFFFF, s)
            self.res.clear(s)
        if lds_builds:
            Lb = mgv.Launch(launch.grid, launch.block, launch.compact,
                            launch.reserved | (mgv.MGV_LAUNCH_FRESH_TABLES if span is not None else 0))
            mgv.call("mgv_batch_run_lds", lb["dq"].vp, lb["di"].vp, lb["dr"].vp, C.byref(lb["plan"]), lb["ws"].vp,
                     lb["ws"].nbytes, Lb, s)
        else:
            mgv.call("mgv_batch_run", bq.vp, bi.vp, bn, self.counter.vp, launch, s)
        mgv.call("mgv_batch_run_lds", lp["dq"].vp, lp["di"].vp, lp["dr"].vp, C.byref(lp["plan"]), lp["ws"].vp,
                 lp["ws"].nbytes, launch, s)

    def capture_ldsb(self, vector_rows: int, launch: mgv.Launch, stream, min_rows: int = 1 << 16,
                     lds_builds: bool = True, fresh: bool = True) -> C.c_void_p:
        self._plan(vector_rows)
        self._ldsb(vector_rows, min_rows)
        if lds_builds:
            self._ldsb(vector_rows, min_rows, builds=True)
        mgv.call("mgv_graph_begin_capture", stream)
        self.issue_ldsb(vector_rows, launch, stream, min_rows, lds_builds, fresh)
        g = C.c_void_p()
        mgv.call("mgv_graph_end_capture", stream, C.byref(g))
        return g

    def items(self, vector_rows: int) -> int:
        (_, _, bn), (_, _, pn) = self._plan(vector_rows)
        return bn + pn

    # -------------------------------------------------------------- plain path
    def issue_plain(self, vector_rows: int, launch: mgv.Launch, stream=None):
        """clear + one launch per vector per query (builds first)."""
        s = stream if stream is not None else self.st
        self.clear(s)
        mgv.ops_call("mgv_ops_run_plain", self.build_q, self.nq, vector_rows, launch, s)
        mgv.ops_call("mgv_ops_run_plain", self.probe_q, self.nq, vector_rows, launch, s)

    # -------------------------------------------------------------- results
    def results(self):
        raw = self.res.download(np.uint64, 8 * 2 * self.nq).reshape(2 * self.nq, 8)
        builds, probes = raw[: self.nq], raw[self.nq:]
        return {"count": probes[:, 0].copy(), "sum": probes[:, 1].view(np.int64).copy(),
                "build_errors": builds[:, 4].copy()}

    def algorithmic_bytes(self, load_factor: float = None) -> int:
        """Probe: 4 B of c per row + 8 B (k, v) per 32-byte granule with a
        survivor (g = 1 - (1 - s)^8) + an 8-byte slot per survivor; build:
        8 B per row read + the table written (tables at load_factor, default
        the batch's)."""
        tot = 0.0
        for x in self.q:
            g = 1.0 - (1.0 - x.sel) ** 8
            tot += x.P * (4 + 8 * g + 8 * x.sel)
            nb = x.nbuckets if load_factor is None else mgv.ht_plan(x.B, load_factor)[0]
            tot += x.B * 8 + nb * 64
        return int(tot)

    def host_columns(self, i: int):
        """Host copies of query i's columns (for the oracle)."""
        x = self.q[i]
        return dict(c=self.rc.download(np.int32, x.P, 4 * x.r_off), k=self.rk.download(np.int32, x.P, 4 * x.r_off),
                    v=self.rv.download(np.int32, x.P, 4 * x.r_off), sk=self.sk.download(np.int32, x.B, 4 * x.s_off),
                    sw=self.sw.download(np.int32, x.B, 4 * x.s_off))

    def close(self):
        for b in (self.rc, self.rk, self.rv, self.sk, self.sw, self.tables, self.res, self.counter):
            b.close()
        self._drop_plans()


# ------------------------------------------------------------------ serving

class ServeTables:
    """Resident base tables of the serving loop.

    Join pairs: for every b in build_logs, S_b has 2^b unique keys (a seeded
    permutation of 0..2^b-1) with payload w in [0, 2^16), and R_b has
    min(16 * 2^b, probe_cap) rows {c uniform in [0, 2^31), k, v in
    [0, 2^16)}.  With join_zipf_s set (the select+join+sum shape of the
    skewed join workload), k follows Zipf(join_zipf_s) over the 2^b build
    keys, ranks mapped through a seeded permutation, so every row matches and
    a few keys take most rows; without it k is uniform in [0, 2^(b+1))
    (about 50% match).
    Group-by tables: G_m has group_rows[m] rows {k ~ Zipf(zipf_s) over
    ngroups ranks mapped through a seeded permutation, v in [0, 2^16)}.
    """

    CFG = 5

    def __init__(self, build_logs=range(16, 27), probe_cap=1 << 29, group_rows=(1 << 30,) * 5,
                 ngroups=1 << 16, zipf_s=0.8, join_zipf_s=None, stream=None):
        sd = mgv.seed
        self.pairs = []
        self.groups = []
        self.ngroups = ngroups
        self.bytes = 0
        for j, b in enumerate(build_logs):
            B = 1 << b
            P = min(16 * B, probe_cap)
            t = {"B": B, "P": P, "log": b}
            for name, rows in (("sk", B), ("sw", B), ("rc", P), ("rk", P), ("rv", P)):
                t[name] = mgv.DeviceBuffer(4 * rows)
                self.bytes += 4 * rows
            mgv.call("mgv_gen_perm", t["sk"].vp, B, B, sd(self.CFG, 0), j, stream)
            mgv.call("mgv_gen_bits", t["sw"].vp, B, sd(self.CFG, 1), j, 16, stream)
            mgv.call("mgv_gen_bits", t["rc"].vp, P, sd(self.CFG, 2), j, 31, stream)
            if join_zipf_s is None:
                mgv.call("mgv_gen_range", t["rk"].vp, P, sd(self.CFG, 3), j, 2 * B, stream)
            else:
                zt = C.c_void_p()
                mgv.call("mgv_zipf_table_create", C.byref(zt), B, float(join_zipf_s))
                mgv.call("mgv_gen_zipf", t["rk"].vp, P, sd(self.CFG, 3), j, zt, B, sd(self.CFG, 8), j, stream)
                mgv.call("mgv_stream_synchronize", stream)
                mgv.call("mgv_free", zt)
            mgv.call("mgv_gen_bits", t["rv"].vp, P, sd(self.CFG, 4), j, 16, stream)
            self.pairs.append(t)
        table = C.c_void_p()
        mgv.call("mgv_zipf_table_create", C.byref(table), ngroups, zipf_s)
        self._zipf = table
        for m, rows in enumerate(group_rows):
            g = {"n": rows, "k": mgv.DeviceBuffer(4 * rows), "v": mgv.DeviceBuffer(4 * rows)}
            self.bytes += 8 * rows
            mgv.call("mgv_gen_zipf", g["k"].vp, rows, sd(self.CFG, 5), m, table, ngroups, sd(self.CFG, 6), m, stream)
            mgv.call("mgv_gen_bits", g["v"].vp, rows, sd(self.CFG, 7), m, 16, stream)
            self.groups.append(g)

    def requests(self, n, rate_qps, seed=0xC5, join_frac=0.7, min_group_log=20, arrivals=True):
        """n requests with Poisson(rate) arrivals (or all at t=0); returns
        (ctypes array of ServeRequest, list of dicts describing them).  The
        queries depend only on `seed`, not on the arrival process."""
        rng = np.random.default_rng(seed)               # the queries
        rng_t = np.random.default_rng(seed ^ 0xA5A5A5)   # the arrival times
        arr = (mgv.ServeRequest * max(1, n))()
        desc = []
        t = 0.0
        for i in range(n):
            if arrivals and rate_qps > 0:
                t += float(rng_t.exponential(1.0 / rate_qps))
            if rng.random() < join_frac:
                j = int(rng.integers(0, len(self.pairs)))
                p = self.pairs[j]
                sel = float(rng.uniform(0.05, 0.95))
                bound = min(INT32_MAX, int(np.floor(sel * 2**31)) - 1)
                arr[i] = mgv.ServeRequest(mgv.MGV_SERVE_JOIN_SUM, bound, p["rc"].ptr, p["rk"].ptr, p["rv"].ptr, p["P"],
                                          p["sk"].ptr, p["sw"].ptr, p["B"], 0, 0, t)
                desc.append({"kind": "join", "pair": j, "bound": bound, "P": p["P"], "B": p["B"], "sel": sel})
            else:
                m = int(rng.integers(0, len(self.groups)))
                g = self.groups[m]
                hi = int(np.log2(g["n"]))
                lo = min(min_group_log, hi)
                rows = int(2 ** rng.uniform(lo, hi)) // 256 * 256 or 256
                rows = min(rows, g["n"])
                arr[i] = mgv.ServeRequest(mgv.MGV_SERVE_GROUPBY, 0, None, g["k"].ptr, g["v"].ptr, rows, None, None, 0,
                                          self.ngroups, 0, t)
                desc.append({"kind": "groupby", "table": m, "rows": rows})
        return arr, desc

    def close(self):
        for t in self.pairs:
            for k in ("sk", "sw", "rc", "rk", "rv"):
                t[k].close()
        for g in self.groups:
            g["k"].close()
            g["v"].close()
        if self._zipf:
            mgv.load().mgv_free(self._zipf)
            self._zipf = None


def serve(requests, n, budget_bytes, max_inflight=4, max_batch=64, vector_rows=65536, load_factor=0.5,
          launch=None):
    """Runs the continuous-batching loop; returns (results array, stats)."""
    cfg = mgv.ServeConfig(budget_bytes, max_inflight, max_batch, vector_rows, 0, load_factor,
                          launch or mgv.launch(0, 256, 1))
    out = (mgv.ServeResult * max(1, n))()
    st = mgv.ServeStats()
    mgv.ops_call("mgv_serve_run", requests, n, C.byref(cfg), out, C.byref(st))
    return out, st


def group_digest(cnt: np.ndarray, s: np.ndarray) -> int:
    """Digest of per-group (COUNT, SUM) arrays, the same chain of splitmix64
    finalisers the serving loop computes on the host."""
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M
        z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M
        return z ^ (z >> 31)

    h = 0x9E3779B97F4A7C15
    for c, x in zip(cnt.astype(np.uint64).tolist(), s.astype(np.int64).tolist()):
        h = mix(h ^ int(c))
        h = mix(h ^ (int(x) & M))
    return h
