"""GPU tests of the completion ring (mgv_ring_*) and of the streaming select
into rotating pool buffers (mgv_select_stream): the device appends each
wave's survivors to the current pool buffer, swaps in a fresh buffer when it
is full, and publishes every full buffer on the ring once all its rows are
written; a host thread copies each buffer out and hands it back with
mgv_pool_release_host while the kernel runs.  The selected multiset must
equal the oracle's with more rows than the pool holds at once, nothing may
be dropped, and every buffer must be back in the pool at the end (the
reference's gpu_to_cpu flow, src/operators/gpu_to_cpu.cu:42-124,149-165,
with buffer_manager::acquire_buffer_blocked_to, src/buffer_manager.cuh:
147-176)."""
from __future__ import annotations

import ctypes as C
import gc
import threading
import time

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu


def make_ring(n):
    r = C.c_void_p()
    mgv.call("mgv_ring_create", n, C.byref(r))
    return r


def poll(ring):
    e = mgv.RingEntry()
    rc = mgv.load().mgv_ring_poll(ring, C.byref(e))
    if rc == 1:
        return None
    mgv.check(rc)
    return e


def test_ring_entries_arrive_in_order_with_results(gpu):
    ring = make_ring(8)
    s = mgv.Stream()
    res = mgv.ResultBuffer(4)
    # results 0..3 carry count = 10 * i + 1, sum = -i
    host = (mgv.Result * 4)()
    for i in range(4):
        host[i].count = 10 * i + 1
        host[i].sum = -i
    mgv.call("mgv_memcpy", res.vp, C.byref(host), C.sizeof(host))
    for i in range(3):
        mgv.call("mgv_ring_emit", ring, 100 + i, 7 * i, res.slot(i), s.vp)
    mgv.call("mgv_ring_emit", ring, 555, 9, None, s.vp)
    s.sync()
    got = [poll(ring) for _ in range(4)]
    assert [(e.id, e.value, e.res.count, e.res.sum) for e in got] == [
        (100, 0, 1, 0), (101, 7, 11, -1), (102, 14, 21, -2), (555, 9, 0, 0)]
    assert poll(ring) is None
    # n results with their ids, published in any order
    ids = mgv.device_array(np.array([7, 3, 9, 1], dtype=np.uint64))
    mgv.call("mgv_ring_emit_results", ring, ids.vp, res.vp, 4, s.vp)
    s.sync()
    got = sorted((e.id, e.value, e.res.count) for e in (poll(ring) for _ in range(4)))
    assert got == [(1, 3, 31), (3, 1, 11), (7, 0, 1), (9, 2, 21)]
    assert poll(ring) is None
    s.close()
    mgv.call("mgv_ring_destroy", ring)


def test_full_ring_is_reported_not_waited_on(gpu):
    ring = make_ring(2)
    s = mgv.Stream()
    for i in range(3):   # the third push finds its slot still taken
        mgv.call("mgv_ring_emit", ring, i, i, None, s.vp)
    s.sync()
    rc = mgv.load().mgv_ring_poll(ring, C.byref(mgv.RingEntry()))
    assert rc == mgv.E_CAPACITY
    s.close()
    mgv.call("mgv_ring_destroy", ring)


class Catcher(threading.Thread):
    """The host half of gpu_to_cpu: takes full buffers off the ring, copies
    their rows out on its own stream, hands the buffers back to the pool.

    While the kernel waits for returned buffers this thread must make no
    call that waits for the whole device (hipFree, hipStreamCreate): its
    stream is created before the launch, and the test turns the garbage
    collector off so that no finalizer of an earlier test's device buffer
    runs here."""

    def __init__(self, pool, ring, rows_per_buffer, host_pool):
        super().__init__(daemon=True)
        self.pool, self.ring, self.cap, self.host_pool = pool, ring, rows_per_buffer, host_pool
        self.parts, self.buffers, self.eof, self.error = [], 0, None, None
        self.max_rows = 0
        self.stream = None if host_pool else mgv.Stream()

    def run(self):
        try:
            s = self.stream
            host = np.empty(self.cap, dtype=np.int32)
            t_end = time.time() + 120
            while time.time() < t_end:
                e = poll(self.ring)
                if e is None:
                    time.sleep(2e-5)
                    continue
                if e.id == mgv.RING_EOF:
                    self.eof = (e.value, e.res.count, e.res.overflow, e.res.reserved[0])
                    break
                assert 0 < e.value <= self.cap, e.value
                data = C.c_void_p()
                mgv.call("mgv_pool_buffer", self.pool, int(e.id), C.byref(data), None)
                if self.host_pool:   # rows already in pinned host memory: read in place
                    rows = (C.c_int32 * e.value).from_address(data.value)
                    self.parts.append(np.ctypeslib.as_array(rows).copy())
                else:
                    mgv.call("mgv_memcpy_async", host.ctypes.data_as(C.c_void_p), data, 4 * e.value, s.vp)
                    s.sync()
                    self.parts.append(host[:e.value].copy())
                self.max_rows = max(self.max_rows, e.value)
                mgv.call("mgv_pool_release_host", self.pool, int(e.id))
                self.buffers += 1
        except Exception as ex:   # reported by the test
            self.error = ex


@pytest.mark.parametrize("n,vector,nbuf,rows,selectivity,host_pool", [
    (1 << 20, 0, 4, 4096, 0.5, True),           # one launch; 2^19 survivors through 16 Ki rows of pool
    (200003, 4096, 3, 1024, 0.9, True),         # per-vector launches as the reference's plan, ragged end
    (1 << 22, 0, 8, 65536, 0.25, True),         # larger buffers
    (5000, 0, 2, 256, 1.0, True),               # every row, buffers of one wave chunk
    (1 << 22, 0, 16, 65536, 0.5, False),        # device-memory pool, copied out by the host
])
def test_select_stream_rotates_more_rows_than_the_pool(gpu, orc, n, vector, nbuf, rows, selectivity, host_pool):
    pool = C.c_void_p()
    mgv.call("mgv_pool_create_host" if host_pool else "mgv_pool_create", nbuf, rows, C.byref(pool))
    ring = make_ring(nbuf + 2)
    so = C.c_void_p()
    mgv.call("mgv_stream_out_create", pool, ring, C.byref(so))
    pred = orc.gen_bits(n, orc.seed(1, 1), 40, 31)
    val = orc.gen_bits(n, orc.seed(1, 0), 40, 20)
    bound = int(selectivity * (1 << 31)) - 1
    dp, dv = mgv.device_array(pred), mgv.device_array(val)
    want = np.sort(val[pred <= bound])
    assert len(want) > nbuf * rows   # more rows than the pool holds at once

    s = mgv.Stream()
    cat = Catcher(pool, ring, rows, host_pool)
    mgv.call("mgv_stream_out_open", so, s.vp)
    s.sync()
    gc.collect()
    gc.disable()
    try:
        cat.start()
        if vector:
            for lo in range(0, n, vector):
                m = min(vector, n - lo)
                mgv.call("mgv_select_stream", dp.at(4 * lo), dv.at(4 * lo), m, bound, so, mgv.launch(0, 256), s.vp)
        else:
            mgv.call("mgv_select_stream", dp.vp, dv.vp, n, bound, so, mgv.launch(), s.vp)
        mgv.call("mgv_stream_out_close", so, s.vp)
        s.sync()
        cat.join(timeout=150)
    finally:
        gc.enable()
    if cat.stream is not None:
        cat.stream.close()
    assert not cat.is_alive(), "catcher did not see the end of the stream"
    assert cat.error is None, cat.error
    assert cat.eof is not None
    eof_rows, rows_handed, dropped, buffers = cat.eof
    assert dropped == 0
    assert eof_rows == rows_handed == len(want)
    assert buffers == cat.buffers
    got = np.sort(np.concatenate(cat.parts)) if cat.parts else np.empty(0, np.int32)
    assert np.array_equal(got, want)
    # conservation: every buffer is back in the pool (stack or return ring)
    avail, prod, stacked = C.c_uint32(), C.c_uint32(), C.c_uint32()
    mgv.call("mgv_pool_stats", pool, C.byref(avail), C.byref(prod), C.byref(stacked), None)
    assert stacked.value == nbuf
    s.close()
    mgv.call("mgv_stream_out_destroy", so)
    mgv.call("mgv_ring_destroy", ring)
    mgv.call("mgv_pool_destroy", pool)
