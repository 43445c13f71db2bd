"""ctypes binding of libmgv.so (the C ABI declared in include/mgv.h).

Used by the tests, bench.py and the autotuner.  Every call goes to the HIP
library; there is no CPU fallback: if libmgv.so is missing or fails to load,
`load()` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmgv.so")

MGV_PAYLOAD_NONE = 0
MGV_PAYLOAD_COLUMN = 1
MGV_PAYLOAD_ROWID = 2
MGV_PROBE_DIGEST = 1
MGV_LAUNCH_FRESH_TABLES = 1 << 14   # mgv_batch_run_lds: LDS-built tables written whole, no clear needed
SLOTS_PER_BUCKET = 8

E_INVALID = 10001
E_NOT_INIT = 10002
E_CAPACITY = 10003
E_ALIGN = 10004
E_RESERVED = 10005
E_NOGPU = 10006


class MgvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mgv error {code}: {msg}")
        self.code = code


class Launch(C.Structure):
    _fields_ = [("grid", C.c_uint32), ("block", C.c_uint32), ("compact", C.c_int32), ("reserved", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("sum", C.c_int64),
        ("digest_sum", C.c_uint64),
        ("digest_xor", C.c_uint64),
        ("errors", C.c_uint64),
        ("overflow", C.c_uint64),
        ("reserved", C.c_uint64 * 2),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("count", "sum", "digest_sum", "digest_xor", "errors", "overflow")}


class RingEntry(C.Structure):
    _fields_ = [("id", C.c_uint64), ("value", C.c_uint64), ("res", Result)]


RING_EOF = (1 << 64) - 1


class DeviceProps(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("arch", C.c_char * 32),
        ("cus", C.c_int32),
        ("wave_size", C.c_int32),
        ("lds_per_block", C.c_int64),
        ("hbm_bytes", C.c_int64),
        ("l2_bytes", C.c_int64),
        ("clock_khz", C.c_int32),
        ("mem_clock_khz", C.c_int32),
        ("mem_bus_width", C.c_int32),
        ("max_threads_per_block", C.c_int32),
    ]


VP = C.c_void_p
U64 = C.c_uint64


MGV_Q_SELECT_SUM = 1
MGV_Q_PROBE_SUM = 2
MGV_Q_PROBE_GROUPBY = 3
MGV_Q_GROUPBY = 4
MGV_Q_BUILD = 5


class Query(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("bound", C.c_int32), ("pred", C.c_void_p), ("key", C.c_void_p),
        ("val", C.c_void_p), ("n", C.c_uint64), ("slots", C.c_void_p), ("nbuckets", C.c_uint32),
        ("ngroups", C.c_uint32), ("gcnt", C.c_void_p), ("gsum", C.c_void_p), ("res", C.c_void_p),
        ("payload_mode", C.c_int32), ("flags", C.c_uint32), ("row_base", C.c_int64),
    ]


class WorkItem(C.Structure):
    _fields_ = [("query", C.c_uint32), ("rows", C.c_uint32), ("lo", C.c_uint64)]


class LdsBatchPlan(C.Structure):
    """mgv_lds_batch_plan: item ranges and workspace of an LDS batch probe."""
    _fields_ = [
        ("nitems", C.c_uint64), ("nscatter", C.c_uint64), ("nparts", C.c_uint64), ("nbuild_parts", C.c_uint64),
        ("build_rows", C.c_uint64), ("region_rows", C.c_uint64), ("workspace_bytes", C.c_uint64),
        ("nqueries", C.c_uint32), ("reserved", C.c_uint32),
    ]


class SlicedPlan(C.Structure):
    """mgv_sliced_plan: item sections and workspace of an XCD-sliced batch."""
    _fields_ = [
        ("nitems", C.c_uint64), ("nscatter", C.c_uint64), ("section", C.c_uint64 * 10),
        ("region_rows", C.c_uint64), ("workspace_bytes", C.c_uint64), ("nqueries", C.c_uint32),
        ("nregions", C.c_uint32),
    ]
U32 = C.c_uint32
I32 = C.c_int32
I64 = C.c_int64

# name -> argtypes (all functions return int status unless listed in _RESTYPE)
SIGNATURES = {
    "mgv_init": [C.c_int],
    "mgv_shutdown": [],
    "mgv_abi_version": [],
    "mgv_last_error_string": [],
    "mgv_get_device_props": [C.POINTER(DeviceProps)],
    "mgv_malloc": [C.POINTER(VP), C.c_size_t],
    "mgv_free": [VP],
    "mgv_host_malloc": [C.POINTER(VP), C.c_size_t],
    "mgv_host_free": [VP],
    "mgv_mem_get_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "mgv_memcpy": [VP, VP, C.c_size_t],
    "mgv_memcpy_async": [VP, VP, C.c_size_t, VP],
    "mgv_fill_u32": [VP, U64, U32, VP],
    "mgv_fill_u64": [VP, U64, U64, VP],
    "mgv_stream_create": [C.POINTER(VP)],
    "mgv_stream_destroy": [VP],
    "mgv_stream_synchronize": [VP],
    "mgv_device_synchronize": [],
    "mgv_event_create": [C.POINTER(VP)],
    "mgv_event_destroy": [VP],
    "mgv_event_record": [VP, VP],
    "mgv_event_synchronize": [VP],
    "mgv_event_query": [VP],
    "mgv_event_elapsed_ms": [C.POINTER(C.c_float), VP, VP],
    "mgv_stream_wait_event": [VP, VP],
    "mgv_graph_begin_capture": [VP],
    "mgv_graph_end_capture": [VP, C.POINTER(VP)],
    "mgv_graph_launch": [VP, VP],
    "mgv_graph_destroy": [VP],
    "mgv_gen_bits": [VP, U64, U64, U32, U32, VP],
    "mgv_gen_range": [VP, U64, U64, U32, U32, VP],
    "mgv_gen_perm": [VP, U64, U32, U64, U32, VP],
    "mgv_gen_iota": [VP, U64, I32, VP],
    "mgv_zipf_table_create": [C.POINTER(VP), U32, C.c_double],
    "mgv_gen_zipf": [VP, U64, U64, U32, VP, U32, U64, U32, VP],
    "mgv_result_clear": [VP, U32, VP],
    "mgv_ht_plan": [U64, C.c_double, C.POINTER(U32), C.POINTER(U64)],
    "mgv_ht_clear": [VP, U32, VP],
    "mgv_hashjoin_build": [VP, U32, VP, VP, C.c_int, VP, I32, U64, I64, VP, Launch, VP],
    "mgv_hashjoin_probe": [VP, U32, VP, I32, VP, VP, U64, I64, U32, VP, Launch, VP],
    "mgv_hashjoin_probe_groupby": [VP, U32, VP, I32, VP, VP, U64, VP, VP, U32, Launch, VP],
    "mgv_hashjoin_probe_materialize": [VP, U32, VP, U64, I64, VP, VP, U64, VP, Launch, VP],
    "mgv_select_sum": [VP, VP, U64, I32, VP, Launch, VP],
    "mgv_select_materialize": [VP, VP, U64, I32, VP, U64, VP, Launch, VP],
    "mgv_groupby": [VP, VP, U64, VP, VP, U32, VP, Launch, VP],
    "mgv_partition_workspace_size": [U64, U32, U32, C.POINTER(U64)],
    "mgv_hashjoin_probe_partitioned": [VP, U32, VP, I32, VP, VP, U64, I64, U32, U32, VP, U64, VP, Launch, VP],
    "mgv_hashjoin_probe_groupby_partitioned": [VP, U32, VP, I32, VP, VP, U64, VP, VP, U32, U32, VP, U64, Launch,
                                               VP],
    "mgv_debug_xcc_histogram": [VP, U32, VP],
    "mgv_batch_plan": [VP, U32, U32, VP, U64, C.POINTER(U64)],
    "mgv_batch_run": [VP, VP, U64, VP, Launch, VP],
    "mgv_batch_plan_sliced": [VP, U32, U32, U64, U32, VP, U64, VP, VP],
    "mgv_batch_run_sliced": [VP, VP, VP, VP, VP, U64, Launch, VP],
    "mgv_batch_plan_lds": [VP, U32, U32, U64, VP, U64, VP, VP],
    "mgv_batch_run_lds": [VP, VP, VP, VP, VP, U64, Launch, VP],
    "mgv_pool_create": [U32, U32, C.POINTER(VP)],
    "mgv_pool_destroy": [VP],
    "mgv_pool_acquire": [VP, C.POINTER(I32), VP],
    "mgv_pool_release": [VP, I32, VP],
    "mgv_pool_register_producer": [VP, VP],
    "mgv_pool_unregister_producer": [VP, VP],
    "mgv_pool_buffer": [VP, I32, C.POINTER(VP), C.POINTER(VP)],
    "mgv_pool_info": [VP, C.POINTER(U32), C.POINTER(U32)],
    "mgv_pool_stats": [VP, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), VP],
    "mgv_pool_stress": [VP, U32, U32, VP, VP],
    "mgv_select_to_buffer": [VP, VP, U64, I32, VP, I32, VP],
    "mgv_channel_create": [U32, C.POINTER(VP)],
    "mgv_channel_destroy": [VP],
    "mgv_channel_emit": [VP, U32, VP, U32, VP],
    "mgv_channel_poll": [VP, U32, VP, C.POINTER(U32)],
    "mgv_ring_create": [U32, C.POINTER(VP)],
    "mgv_ring_destroy": [VP],
    "mgv_ring_emit": [VP, U64, U64, VP, VP],
    "mgv_ring_emit_results": [VP, VP, VP, U32, VP],
    "mgv_ring_poll": [VP, VP],
    "mgv_pool_release_host": [VP, I32],
    "mgv_pool_create_host": [U32, U32, C.POINTER(VP)],
    "mgv_stream_out_create": [VP, VP, C.POINTER(VP)],
    "mgv_stream_out_destroy": [VP],
    "mgv_stream_out_open": [VP, VP],
    "mgv_stream_out_close": [VP, VP],
    "mgv_select_stream": [VP, VP, U64, I32, VP, Launch, VP],
    "mgv_groupby_workspace_size": [U64, U32, C.POINTER(U64)],
    "mgv_probe_lds_workspace_size": [U64, U32, U32, C.POINTER(U64)],
    "mgv_build_lds_workspace_size": [U64, U32, U32, C.POINTER(U64)],
    "mgv_hashjoin_build_lds": [VP, U32, VP, VP, I32, VP, I32, U64, I64, U32, VP, U64, VP, Launch, VP],
    "mgv_hashjoin_probe_lds": [VP, U32, VP, I32, VP, VP, U64, U32, VP, U64, VP, Launch, VP],
    "mgv_hashjoin_probe_lds_noalloc": [VP, U32, VP, I32, VP, VP, U64, U32, VP, Launch, VP],
    "mgv_probe_groupby_lds_workspace_size": [U64, U32, U32, C.POINTER(U64)],
    "mgv_hashjoin_probe_groupby_lds": [VP, U32, VP, I32, VP, VP, U64, VP, VP, U32, U32, VP, U64, Launch, VP],
    "mgv_groupby_partitioned": [VP, VP, U64, VP, VP, U32, VP, U64, VP, Launch, VP],
    "mgv_split": [VP, I32, VP, U32, U64, U32, VP, U64, VP, Launch, VP],
}
POOL_END = -1
POOL_RETRY = -2
_RESTYPE = {"mgv_last_error_string": C.c_char_p}

_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libmgv.so and declare the signatures; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"libmgv.so not built at {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


# ------------------------------------------------------------ libmgv_ops.so

MGV_SERVE_JOIN_SUM = 1
MGV_SERVE_GROUPBY = 2


class ServeRequest(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("bound", C.c_int32), ("pred", C.c_void_p), ("key", C.c_void_p),
        ("val", C.c_void_p), ("n", C.c_uint64), ("bkey", C.c_void_p), ("bval", C.c_void_p), ("bn", C.c_uint64),
        ("ngroups", C.c_uint32), ("reserved", C.c_uint32), ("arrival_s", C.c_double),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("count", C.c_uint64), ("sum", C.c_int64), ("digest", C.c_uint64), ("footprint", C.c_uint64),
        ("admit_s", C.c_double), ("done_s", C.c_double), ("latency_s", C.c_double), ("batch", C.c_uint32),
        ("errors", C.c_uint32),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32), ("max_batch", C.c_uint32),
        ("vector_rows", C.c_uint32), ("reserved", C.c_uint32), ("load_factor", C.c_double), ("launch", Launch),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("elapsed_s", C.c_double), ("queries_per_s", C.c_double), ("p50_s", C.c_double), ("p99_s", C.c_double),
        ("max_s", C.c_double), ("mean_s", C.c_double), ("batches", C.c_uint32), ("max_inflight_seen", C.c_uint32),
        ("peak_bytes", C.c_uint64), ("mean_batch", C.c_double), ("ring_results", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


OPS_PATH = os.path.join(HERE, "lib", "libmgv_ops.so")
OPS_SIGNATURES = {
    "mgv_ops_init": [U32, U32, U32],
    "mgv_ops_shutdown": [],
    "mgv_ops_pool_stats": [C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)],
    "mgv_ops_last_error": [],
    "mgv_ops_sum_select": [VP, VP, U64, I32, U32, C.c_int, U32, U32, VP, C.POINTER(I64)],
    "mgv_ops_select": [VP, VP, U64, I32, C.c_int, U32, C.c_int, U32, U32, VP, U64, C.POINTER(U64)],
    "mgv_ops_select_split": [VP, VP, U64, I32, U32, C.c_int, U32, U32, VP, U64, C.POINTER(U64)],
    "mgv_ops_split_sums": [VP, VP, U64, I32, U32, C.c_int, U32, U32, VP, VP],
    "mgv_ops_sum_select_streams": [VP, VP, U64, I32, U32, C.c_int, U32, U32, U32, C.POINTER(I64)],
    "mgv_ops_select_streams": [VP, VP, U64, I32, U32, C.c_int, U32, U32, U32, VP, U64, C.POINTER(U64)],
    "mgv_ops_select_pipelines": [VP, VP, U64, I32, U32, C.c_int, U32, U32, U32, C.c_int, VP, U64, C.POINTER(I64)],
    "mgv_ops_count_join": [VP, VP, U64, VP, VP, U64, I32, U32, C.c_int, U32, U32, VP, C.POINTER(U64)],
    "mgv_ops_run_plain": [VP, U32, U32, Launch, VP],
    "mgv_serve_run": [VP, U32, VP, VP, VP],
}
_ops: Optional[C.CDLL] = None


def load_ops(path: str = OPS_PATH) -> C.CDLL:
    """Load libmgv_ops.so (the C++ operator layer); raises if it is missing."""
    global _ops
    if _ops is not None:
        return _ops
    load()
    if not os.path.exists(path):
        raise FileNotFoundError(f"libmgv_ops.so not built at {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, args in OPS_SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "mgv_ops_last_error" else C.c_int
    _ops = lib
    return lib


def ops_call(name: str, *args) -> None:
    rc = getattr(load_ops(), name)(*args)
    if rc != 0:
        raise MgvError(rc, _ops.mgv_ops_last_error().decode())


def check(rc: int) -> None:
    if rc != 0:
        msg = _lib.mgv_last_error_string().decode() if _lib else "?"
        raise MgvError(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def launch(grid: int = 0, block: int = 256, compact: int = -1) -> Launch:
    return Launch(grid, block, compact, 0)


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    call("mgv_init", device)
    _initialised = True


def props() -> DeviceProps:
    p = DeviceProps()
    call("mgv_get_device_props", C.byref(p))
    return p


class DeviceBuffer:
    """Device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = VP()
        call("mgv_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @property
    def vp(self) -> VP:
        # The pointer object keeps this buffer alive while it is referenced
        # (e.g. inside an argument list), so `device_array(x).vp` cannot be
        # freed before the call that uses it.
        v = VP(self.ptr)
        v._owner = self
        return v

    def close(self):
        if self.ptr:
            load().mgv_free(VP(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, arr: np.ndarray, offset: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        call("mgv_memcpy", VP(self.ptr + offset), a.ctypes.data_as(VP), a.nbytes)
        return self

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        if out.nbytes:
            call("mgv_memcpy", out.ctypes.data_as(VP), VP(self.ptr + offset), out.nbytes)
        return out

    def at(self, byte_offset: int) -> VP:
        v = VP(self.ptr + byte_offset)
        v._owner = self
        return v


def device_array(arr: np.ndarray) -> DeviceBuffer:
    a = np.ascontiguousarray(arr)
    return DeviceBuffer(a.nbytes).upload(a)


class ResultBuffer(DeviceBuffer):
    """Array of n mgv_result on the device."""

    def __init__(self, n: int = 1):
        super().__init__(C.sizeof(Result) * n)
        self.n = n
        self.clear()

    def clear(self, stream=None):
        call("mgv_result_clear", self.vp, self.n, stream)

    def read(self, i: int = 0) -> Result:
        r = Result()
        call("mgv_memcpy", C.byref(r), VP(self.ptr + i * C.sizeof(Result)), C.sizeof(Result))
        return r

    def slot(self, i: int) -> VP:
        v = VP(self.ptr + i * C.sizeof(Result))
        v._owner = self
        return v


def ht_plan(nkeys: int, load_factor: float = 0.5):
    nb = U32()
    nbytes = U64()
    call("mgv_ht_plan", nkeys, load_factor, C.byref(nb), C.byref(nbytes))
    return nb.value, nbytes.value


class HashTable:
    def __init__(self, nkeys: int, load_factor: float = 0.5):
        self.nbuckets, nbytes = ht_plan(nkeys, load_factor)
        self.buf = DeviceBuffer(nbytes)
        self.clear()

    def clear(self, stream=None):
        call("mgv_ht_clear", self.buf.vp, self.nbuckets, stream)

    def slots(self) -> np.ndarray:
        return self.buf.download(np.uint64, self.nbuckets * SLOTS_PER_BUCKET)


def partition_workspace(n: int, fanout: int, ncols: int = 2) -> "DeviceBuffer":
    b = U64()
    call("mgv_partition_workspace_size", n, fanout, ncols, C.byref(b))
    return DeviceBuffer(b.value)


class Stream:
    def __init__(self):
        p = VP()
        call("mgv_stream_create", C.byref(p))
        self.ptr = p.value

    @property
    def vp(self):
        return VP(self.ptr)

    def sync(self):
        call("mgv_stream_synchronize", self.vp)

    def close(self):
        if self.ptr:
            load().mgv_stream_destroy(VP(self.ptr))
            self.ptr = None


class Event:
    def __init__(self):
        p = VP()
        call("mgv_event_create", C.byref(p))
        self.ptr = p.value

    def record(self, stream=None):
        call("mgv_event_record", VP(self.ptr), stream)

    def sync(self):
        call("mgv_event_synchronize", VP(self.ptr))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        call("mgv_event_elapsed_ms", C.byref(ms), VP(self.ptr), VP(later.ptr))
        return ms.value

    def close(self):
        if self.ptr:
            load().mgv_event_destroy(VP(self.ptr))
            self.ptr = None


def sync():
    call("mgv_device_synchronize")


def seed(config: int, column: int) -> int:
    """Seed of (config, column): 0x4D47560000000000 | config<<8 | column."""
    return 0x4D47560000000000 | (config << 8) | column
