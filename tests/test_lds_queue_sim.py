"""CPU regression test of the per-wave continuation queue of k_lds_probe
(multigpu-vector_amd/csrc/mgv_probe_lds.hip).

A 64-lane simulation of one wave runs the kernel's row loop, lds_probe_quad,
queue_push and queue_round on a small table slice, with the same queue size
(128 entries), the same wave-uniform head/tail counters, the same
LDS-or-global bucket choice and the same chain end rule.  Every queue read
must see an entry that was written and not yet consumed, every bucket index
must lie in the slice or the table, the queue never holds 128 or more
entries, and COUNT / SUM must equal a brute-force join.

The kernel once let each lane leave the row loop on its own; the lanes that
left kept stale head/tail copies and the final drain read unwritten queue
slots as bucket indices (an out-of-bounds table read).  The same simulation
with per-lane counters is run too and must trip the invariant checks: that
is what the wave-uniform loop (while ballot(i0 < r1)) and the readfirstlane
counters prevent.
"""
from __future__ import annotations

import numpy as np
import pytest

WAVE = 64
KQ = 128          # kQueue
EXTRA = 8         # kSliceExtra
EMPTY = 0xFFFFFFFF
M32 = 0xFFFFFFFF


def fmix32(h):
    h &= M32
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    h ^= h >> 16
    return h


def bucket_of_hash(h, nb):
    return (h * nb) >> 32


class QueueFault(AssertionError):
    pass


def build_table(keys, pays, nb):
    """Open addressing with 8-slot buckets filled in slot order, chains
    continuing into the next bucket (wrapping), as the library's build."""
    tab = [[(EMPTY, EMPTY)] * 8 for _ in range(nb)]
    for k, p in zip(keys, pays):
        b = bucket_of_hash(fmix32(k), nb)
        while True:
            row = tab[b]
            free = [s for s in range(8) if row[s] == (EMPTY, EMPTY)]
            if free:
                row[free[0]] = (k & M32, p & M32)
                break
            b = (b + 1) % nb
    return tab


class Wave:
    """One wave of k_lds_probe working on the slice [lo, lo + slice)."""

    def __init__(self, tab, nb, lo, slice_buckets, uniform):
        self.tab, self.nb, self.lo = tab, nb, lo
        self.nload = min(slice_buckets + EXTRA, nb - lo)
        self.uniform = uniform
        # LDS slice: hashed keys, payloads, fill count
        self.skey, self.spay, self.scnt = [], [], []
        for i in range(self.nload):
            row = tab[lo + i]
            c = next((s for s in range(8) if row[s] == (EMPTY, EMPTY)), 8)
            self.skey.append([fmix32(row[s][0]) for s in range(8)])
            self.spay.append([row[s][1] for s in range(8)])
            self.scnt.append(c)
        self.qkey = [None] * KQ
        self.qval = [None] * KQ
        self.qbkt = [None] * KQ
        self.live = [False] * KQ          # written and not yet consumed
        self.head = [0] * WAVE            # per-lane copies of the counters
        self.tail = [0] * WAVE
        self.cnt = 0
        self.sum = 0
        self.max_fill = 0

    # ---------------------------------------------------------- helpers
    def _lane_counters(self, lanes):
        hs = {self.head[l] for l in lanes}
        ts = {self.tail[l] for l in lanes}
        if self.uniform:
            assert len(hs) == 1 and len(ts) == 1
        return hs, ts

    def _slice_bucket(self, idx, h):
        assert 0 <= idx < self.nload, "LDS slice index out of range"
        c = self.scnt[idx]
        hm = [s for s in range(c) if self.skey[idx][s] == h]
        return len(hm), sum(int(np.int32(np.uint32(self.spay[idx][s]))) for s in hm), c == 8

    def _global_bucket(self, b, h):
        if not 0 <= b < self.nb:
            raise QueueFault(f"table read at bucket {b} of {self.nb}")
        row = self.tab[b]
        hits = [s for s in range(8) if row[s] != (EMPTY, EMPTY) and fmix32(row[s][0]) == h]
        full = row[7] != (EMPTY, EMPTY)
        return len(hits), sum(int(np.int32(np.uint32(row[s][1]))) for s in hits), full

    def push(self, lanes, want, key, val, bkt):
        """queue_push over the active lanes: ballot, mbcnt slot, new tail."""
        rank = 0
        for l in lanes:
            if want[l]:
                pos = (self.tail[l] + rank) & (KQ - 1)
                if self.live[pos]:
                    raise QueueFault("queue overwrite of an unconsumed entry")
                self.qkey[pos], self.qval[pos], self.qbkt[pos] = key[l], val[l], bkt[l]
                self.live[pos] = True
                rank += 1
        for l in lanes:
            self.tail[l] += rank
        fill = max(self.tail[l] - self.head[l] for l in lanes)
        self.max_fill = max(self.max_fill, fill)
        if fill >= KQ:
            raise QueueFault("queue holds more than it can")

    def round(self, lanes):
        """queue_round: up to 64 queued chains, one per lane."""
        key, val, bkt, mine = {}, {}, {}, {}
        for l in lanes:
            take = min(self.tail[l] - self.head[l], WAVE)
            mine[l] = l < take
            if mine[l]:
                pos = (self.head[l] + l) & (KQ - 1)
                if not self.live[pos] or self.qbkt[pos] is None:
                    raise QueueFault(f"lane {l} read queue slot {pos}, which holds no entry")
                key[l], val[l], bkt[l] = self.qkey[pos], self.qval[pos], self.qbkt[pos]
        for l in lanes:
            if mine[l]:
                self.live[(self.head[l] + l) & (KQ - 1)] = False
        more, nkey, nval, nbkt = {}, {}, {}, {}
        for l in lanes:
            more[l] = False
            if not mine[l]:
                continue
            idx = bkt[l] - self.lo
            if 0 <= idx < self.nload:
                hits, psum, full = self._slice_bucket(idx, key[l])
            else:
                hits, psum, full = self._global_bucket(bkt[l], key[l])
            self.cnt += hits
            self.sum += psum + hits * val[l]
            nxt = 0 if bkt[l] + 1 == self.nb else bkt[l] + 1
            more[l] = full and nxt != bucket_of_hash(key[l], self.nb)
            nkey[l], nval[l], nbkt[l] = key[l], val[l], nxt
        for l in lanes:
            self.head[l] += min(self.tail[l] - self.head[l], WAVE)
        self.push(lanes, more, nkey, nval, nbkt)

    def probe_quad(self, lanes, k, v, m):
        """lds_probe_quad for the active lanes: slice lookups, then one
        push per row j followed by rounds while a full wave is queued."""
        for j in range(4):
            go, key, val, frm = {}, {}, {}, {}
            for l in lanes:
                h = k[l][j]
                b = bucket_of_hash(h, self.nb)
                idx = b - self.lo
                live = bool((m[l] >> j) & 1)
                ins = live and 0 <= idx < self.nload
                c = 0
                if ins:
                    hits, psum, full = self._slice_bucket(idx, h)
                    c = 8 if full else 0
                    self.cnt += hits
                    self.sum += psum + hits * v[l][j]
                nxt = (0 if b + 1 == self.nb else b + 1) if ins else b
                go[l] = live and (not ins or (c == 8 and nxt != b))
                key[l], val[l], frm[l] = h, v[l][j], nxt
            self.push(lanes, go, key, val, frm)
            while True:
                self._lane_counters(lanes)
                if not any(self.tail[l] - self.head[l] >= WAVE for l in lanes):
                    break
                self.round(lanes)

    def run(self, hashes, vals, stride):
        """The row loop of one wave over rows [0, n) and the final drain."""
        n = len(hashes)
        i0 = [4 * l for l in range(WAVE)]

        def quad(i):
            m = sum(1 << j for j in range(4) if i + j < n)
            k = [int(hashes[i + j]) if i + j < n else 0 for j in range(4)]
            v = [int(vals[i + j]) if i + j < n else 0 for j in range(4)]
            return k, v, m

        while True:
            if self.uniform:
                # while (ballot(i0 < r1)): every lane iterates, finished lanes
                # carry an empty mask
                if not any(i < n for i in i0):
                    break
                lanes = list(range(WAVE))
            else:
                # the faulty form: a lane leaves as soon as its own rows end
                lanes = [l for l in range(WAVE) if i0[l] < n]
                if not lanes:
                    break
            k, v, m = {}, {}, {}
            for l in lanes:
                k[l], v[l], m[l] = quad(i0[l])
            self.probe_quad(lanes, k, v, m)
            for l in lanes:
                i0[l] += stride
        # drain: while (tail != head) head = queue_round(...)
        all_lanes = list(range(WAVE))
        while any(self.tail[l] != self.head[l] for l in all_lanes):
            self.round([l for l in all_lanes if self.tail[l] != self.head[l]] if not self.uniform else all_lanes)
        return self.cnt, self.sum


def brute(keys, pays, rk, rv):
    from collections import defaultdict
    d = defaultdict(list)
    for k, p in zip(keys, pays):
        d[int(k)].append(int(p))
    cnt = tot = 0
    for k, v in zip(rk, rv):
        ps = d.get(int(k), [])
        cnt += len(ps)
        tot += sum(ps) + len(ps) * int(v)
    return cnt, tot


def make_case(seed, nb, nkeys, distinct, slice_buckets, lo, nrows, outside):
    rng = np.random.default_rng(seed)
    keys = rng.integers(-(1 << 31), 1 << 31, distinct, dtype=np.int64)
    keys = rng.choice(keys, nkeys).astype(np.int64)              # repeated keys
    pays = rng.integers(-(1 << 15), 1 << 16, nkeys)
    tab = build_table([int(k) for k in keys], [int(p) for p in pays], nb)
    # probe rows: keys whose home bucket lies in the slice (as after the
    # level-2 partitioning), some build keys, some misses, some homes
    # outside the slice
    pool = [int(k) for k in keys] + [int(x) for x in rng.integers(-(1 << 31), 1 << 31, 4 * distinct)]
    homes = {k: bucket_of_hash(fmix32(k), nb) for k in pool}
    inside = [k for k in pool if lo <= homes[k] < lo + slice_buckets]
    other = [k for k in pool if not lo <= homes[k] < lo + slice_buckets]
    rk = list(rng.choice(inside, nrows))
    for i in rng.choice(nrows, outside, replace=False):
        rk[i] = int(rng.choice(other))
    rv = rng.integers(0, 1 << 16, nrows)
    return keys, pays, tab, np.array(rk, dtype=np.int64), rv


CASES = [
    # seed, nb, keys, distinct, slice, lo, rows, rows with a home outside
    (1, 64, 460, 300, 16, 16, 3000, 40),      # load factor 0.9: chains leave the slice
    (2, 64, 500, 60, 16, 48, 2500, 0),        # heavy duplicates, last slice: chains wrap to bucket 0
    (3, 32, 250, 250, 32, 0, 777, 0),         # one slice is the whole table
    (4, 128, 512, 400, 32, 64, 130, 5),       # fewer rows than a wave's first quads
]


@pytest.mark.parametrize("case", CASES)
def test_wave_uniform_queue_keeps_its_invariants(case):
    seed, nb, nkeys, distinct, sl, lo, nrows, outside = case
    keys, pays, tab, rk, rv = make_case(seed, nb, nkeys, distinct, sl, lo, nrows, outside)
    w = Wave(tab, nb, lo, sl, uniform=True)
    got = w.run([fmix32(int(k)) for k in rk], rv, stride=4 * WAVE)
    assert got == brute(keys, pays, rk, rv)
    assert w.max_fill < KQ
    assert not any(w.live)


def test_per_lane_loop_exit_reads_stale_queue_slots():
    """The faulty form (lanes leave the row loop one by one, counters kept
    per lane) makes the drain work from stale counters: it reads or
    overwrites queue slots that are not its own."""
    seed, nb, nkeys, distinct, sl, lo, nrows, outside = CASES[0]
    keys, pays, tab, rk, rv = make_case(seed, nb, nkeys, distinct, sl, lo, nrows + 37, outside)
    w = Wave(tab, nb, lo, sl, uniform=False)
    with pytest.raises(QueueFault):
        w.run([fmix32(int(k)) for k in rk], rv, stride=4 * WAVE)
