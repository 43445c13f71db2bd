"""CPU model check of the rotating-output protocol of mgv_select_stream
(multigpu-vector_amd/csrc/mgv_pool.hpp, mgv_stream.hip): waves reserve rows
in the current pool buffer with a CAS on its {cnt, gen} word, the first
reservation that does not fit seals the buffer, the last of the sealer and
the writers hands it to the host ring, a wave swaps in a fresh buffer (its
generation raised) with a CAS of the {gen, index} output word, and a host
thread copies buffers out and hands them back through the return ring.

Every shared-memory step of a wave is a scheduling point, and a seeded
random scheduler interleaves many waves and the host.  Checked on every
run: a buffer reaches the host only with all of its rows written; every row
reaches the host exactly once (none is dropped while buffers keep coming
back); no reservation ever lands in a buffer that sits in the pool; at the
end every buffer is back in the pool.  The same model with the generation
check taken out (the reference's plain atomicAdd reservation and untagged
CAS, src/buffer.cuh:92-105, src/buffer_manager.cuh:147-176) must be caught
losing or misplacing rows, which shows the checks have teeth."""
from __future__ import annotations

import random

import pytest

NULL = 0xFFFFFFFF
SEAL = 1 << 31
STALE, BUSY = 0xFFFFFFFE, 0xFFFFFFFD


class Model:
    def __init__(self, nbuf, cap, seed, tagged=True):
        self.cap = cap
        self.tagged = tagged
        self.rng = random.Random(seed)
        self.cnt = [0] * nbuf
        self.gen = [0] * nbuf
        self.done = [0] * nbuf
        self.final = [0] * nbuf
        self.data = [[None] * cap for _ in range(nbuf)]
        self.where = ["pool"] * nbuf          # pool | out (current or retiring) | host
        self.stack = list(range(nbuf))
        self.ret = []                          # host -> device return ring
        self.cur = NULL                        # after k_stream_open
        self.ring = []                         # device -> host entries (buffer, rows)
        self.host_rows = []
        self.dropped = 0
        self.errors = []

    # ---------------------------------------------------------------- device steps

    def pool_acquire(self):
        for _ in range(100):
            yield
            if self.stack:
                return self.stack.pop()
            yield
            if self.ret:
                return self.ret.pop(0)
        return None

    def retire(self, b, rows):
        if any(x is None for x in self.data[b][:rows]):
            self.errors.append(f"buffer {b} handed on with unwritten rows")
        if rows == 0:
            self.where[b] = "pool"
            self.stack.append(b)
        else:
            self.where[b] = "host"
            self.ring.append((b, rows, list(self.data[b][:rows])))

    def seal(self, b, v):
        self.final[b] = v
        yield
        old = self.done[b]
        self.done[b] = old + SEAL - v
        if old + SEAL - v == SEAL:
            self.retire(b, v)

    def written(self, b, n):
        old = self.done[b]
        self.done[b] = old + n
        if old + n == SEAL:
            self.retire(b, self.final[b])

    def reserve(self, b, w, n):
        s = (self.gen[b], self.cnt[b])
        yield
        for _ in range(256):
            if self.tagged and s[0] != (w >> 32):
                return STALE
            if s[1] > self.cap:
                return NULL
            if not self.tagged:                       # fetch_add: lands whatever the buffer's state
                old = self.cnt[b]
                self.cnt[b] = old + n
                if self.where[b] == "pool":
                    self.errors.append(f"reservation in buffer {b} while it sits in the pool")
                return old
            if (self.gen[b], self.cnt[b]) == s:       # CAS succeeds
                self.cnt[b] = s[1] + n
                if self.where[b] == "pool":
                    self.errors.append(f"reservation in buffer {b} while it sits in the pool")
                return s[1]
            s = (self.gen[b], self.cnt[b])
            yield
        return BUSY

    def swap(self, seen):
        yield
        if self.cur != seen:
            return self.cur
        nb = yield from self.pool_acquire()
        if nb is None:
            return (seen & ~0xFFFFFFFF) | NULL
        tag = self.gen[nb] + 1 if self.tagged else 0
        self.done[nb] = 0
        self.final[nb] = 0
        yield
        self.gen[nb], self.cnt[nb] = tag, 0
        self.data[nb] = [None] * self.cap
        yield
        nw = (tag << 32) | nb
        if self.cur == seen:
            self.cur = nw
            self.where[nb] = "out"
            return nw
        self.stack.append(nb)
        return self.cur

    def append(self, rows, attempts=4000):
        n = len(rows)
        tried = None
        attempt = 0
        while True:
            yield
            w = self.cur
            b = w & 0xFFFFFFFF
            if b != NULL and w != tried:
                off = yield from self.reserve(b, w, n)
                if off == BUSY:
                    attempt += 1
                    continue
                if off < BUSY and off + n <= self.cap:
                    for i, r in enumerate(rows):
                        if self.data[b][off + i] is not None:
                            self.errors.append(f"slot {off + i} of buffer {b} written twice")
                        self.data[b][off + i] = r
                        if i % 64 == 0:
                            yield
                    yield
                    self.written(b, n)
                    return
                if off <= self.cap:
                    yield from self.seal(b, off)
                tried = w
            r = yield from self.swap(w)
            if r & 0xFFFFFFFF == NULL:
                attempt += 1
                if attempt >= attempts:
                    self.dropped += n
                    return

    def wave(self, chunks):
        for rows in chunks:
            yield from self.append(rows)

    def close(self):
        w = self.cur
        b = w & 0xFFFFFFFF
        if b != NULL:
            off = BUSY
            while off == BUSY:
                off = yield from self.reserve(b, w, self.cap + 1)
            if off <= self.cap:
                yield from self.seal(b, off)
            self.cur = NULL

    # ---------------------------------------------------------------- host

    def host_step(self):
        if self.ring:
            b, rows, snap = self.ring.pop(0)
            if snap != self.data[b][:rows]:
                self.errors.append(f"buffer {b} changed after it was handed on")
            self.host_rows.extend(snap)
            self.where[b] = "pool"
            self.ret.append(b)


def run(nbuf, cap, nwaves, rows_per_wave, seed, tagged=True, host_speed=0.3, max_steps=5_000_000):
    """Runs the model to the end; returns it, the appended rows and whether
    every wave finished within max_steps scheduling steps."""
    m = Model(nbuf, cap, seed, tagged)
    rng = random.Random(seed + 1)
    expected = []
    waves = []
    for w in range(nwaves):
        chunks, i = [], 0
        while i < rows_per_wave:
            k = min(rng.choice([256, 256, 256, rng.randint(1, 256)]), rows_per_wave - i)
            chunks.append([(w, i + j) for j in range(k)])
            expected.extend(chunks[-1])
            i += k
        waves.append(m.wave(chunks))
    live = list(waves)
    steps = 0
    while live:
        steps += 1
        if steps >= max_steps:
            return m, expected, False
        if rng.random() < host_speed:
            m.host_step()
        g = rng.choice(live)
        try:
            next(g)
        except StopIteration:
            live.remove(g)
    closer = m.close()
    for _ in closer:
        pass
    while m.ring:
        m.host_step()
    return m, expected, True


@pytest.mark.parametrize("seed", range(12))
def test_protocol_delivers_every_row_once(seed):
    nbuf = [2, 3, 4, 8][seed % 4]
    cap = [256, 512, 1024][seed % 3]
    m, expected, finished = run(nbuf, cap, nwaves=6 + seed % 5, rows_per_wave=700 + 97 * seed, seed=seed)
    assert finished
    assert not m.errors, m.errors[:5]
    assert m.dropped == 0
    assert sorted(m.host_rows) == sorted(expected)
    # conservation: every buffer is back in the pool (stack or return ring)
    assert sorted(m.stack + m.ret) == list(range(nbuf))
    assert all(w == "pool" for w in m.where)


def test_slow_host_drops_and_reports_but_never_loses():
    # a host that almost never hands buffers back: waves give up after their
    # attempts and count the rows as dropped; what did reach the host is
    # still exact and complete
    m, expected, finished = run(2, 256, nwaves=4, rows_per_wave=1500, seed=99, host_speed=0.0005)
    assert finished
    assert not m.errors, m.errors[:5]
    assert len(m.host_rows) + m.dropped == len(expected)
    assert len(set(m.host_rows)) == len(m.host_rows)
    assert set(m.host_rows) <= set(expected)


def test_untagged_reservations_are_caught():
    # the same protocol with a plain fetch_add reservation and untagged
    # words: the model must find rows lost, duplicated or written into a
    # pooled buffer (or a buffer that never reaches the host, which leaves
    # waves retrying) in at least one of these schedules
    bad = 0
    for seed in range(40):
        m, expected, finished = run(2, 256, nwaves=8, rows_per_wave=600, seed=1000 + seed, tagged=False,
                                    host_speed=0.9, max_steps=400_000)
        if not finished or m.errors or sorted(m.host_rows) != sorted(expected):
            bad += 1
    assert bad > 0
