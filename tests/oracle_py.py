"""numpy-friendly ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure: loaded only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

VP, U64, U32, I32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32


class JoinResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("sum", C.c_int64), ("digest_sum", C.c_uint64),
                ("digest_xor", C.c_uint64), ("count32", C.c_int32)]


class ChainedTable(C.Structure):
    _fields_ = [("heads", VP), ("next", VP), ("keys", VP), ("payload", VP), ("log_buckets", C.c_uint32),
                ("n", C.c_uint64)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(VP)


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            subprocess.run(["make", "-C", os.path.dirname(path)], check=True, capture_output=True)
        L = C.CDLL(path)
        L.orc_philox4x32_10.argtypes = [C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]
        L.orc_seed.argtypes = [U32, U32]
        L.orc_seed.restype = U64
        L.orc_word.argtypes = [U64, U32, U64]
        L.orc_word.restype = U32
        L.orc_gen_bits.argtypes = [VP, U64, U64, U32, U32]
        L.orc_gen_range.argtypes = [VP, U64, U64, U32, U32]
        L.orc_feistel.argtypes = [U32, U32, U64, U32]
        L.orc_feistel.restype = U32
        L.orc_perm.argtypes = [U32, U32, U64, U32]
        L.orc_perm.restype = U32
        L.orc_gen_perm.argtypes = [VP, U64, U32, U64, U32]
        L.orc_zipf_cdf.argtypes = [VP, U32, C.c_double]
        L.orc_zipf_rank.argtypes = [U32, VP, U32]
        L.orc_zipf_rank.restype = U32
        L.orc_gen_zipf.argtypes = [VP, U64, U64, U32, VP, U32, U64, U32]
        L.orc_fmix32.argtypes = [U32]
        L.orc_fmix32.restype = U32
        L.orc_mix64.argtypes = [U64]
        L.orc_mix64.restype = U64
        L.orc_split.argtypes = [VP, I32, VP, U32, U64, U32, VP, VP]
        L.orc_select.argtypes = [VP, U64, I32, VP]
        L.orc_select.restype = U64
        L.orc_select_sum.argtypes = [VP, VP, U64, I32, C.POINTER(U64), C.POINTER(C.c_int64), C.POINTER(I32)]
        L.orc_select_sum_mt.argtypes = [VP, VP, U64, I32, C.POINTER(U64), C.POINTER(C.c_int64), C.c_int]
        L.orc_ct_build.argtypes = [C.POINTER(ChainedTable), VP, VP, U64, U32]
        L.orc_ct_free.argtypes = [C.POINTER(ChainedTable)]
        L.orc_ct_build_mt.argtypes = [C.POINTER(ChainedTable), VP, VP, U64, U32, C.c_int]
        L.orc_join_dense.argtypes = [VP, VP, U64, U32, VP, I32, VP, VP, U64, C.POINTER(U64), C.POINTER(C.c_int64),
                                     C.c_int]
        L.orc_ct_probe_groupby_mt.argtypes = [C.POINTER(ChainedTable), VP, I32, VP, VP, U64, VP, VP, U32, C.c_int]
        L.orc_ct_probe.argtypes = [C.POINTER(ChainedTable), VP, I32, VP, VP, U64, U64, C.POINTER(JoinResult)]
        L.orc_ct_probe_mt.argtypes = [C.POINTER(ChainedTable), VP, I32, VP, VP, U64, C.POINTER(JoinResult), C.c_int]
        L.orc_ct_probe_groupby.argtypes = [C.POINTER(ChainedTable), VP, I32, VP, VP, U64, VP, VP, U32]
        L.orc_groupby.argtypes = [VP, VP, U64, VP, VP, U32]
        L.orc_selfjoin_count.argtypes = [VP, U64, I32, U32]
        L.orc_selfjoin_count.restype = C.c_int64
        self.L = L

    # -------------------------------------------------------------- datagen
    def philox(self, ctr, key):
        c = (U32 * 4)(*ctr)
        k = (U32 * 2)(*key)
        o = (U32 * 4)()
        self.L.orc_philox4x32_10(c, k, o)
        return list(o)

    def seed(self, config, column):
        return self.L.orc_seed(config, column)

    def word(self, seed, stream, i):
        return self.L.orc_word(seed, stream, i)

    def gen_bits(self, n, seed, stream, bits):
        a = np.empty(n, np.int32)
        self.L.orc_gen_bits(_ptr(a), n, seed, stream, bits)
        return a

    def gen_range(self, n, seed, stream, rng):
        a = np.empty(n, np.int32)
        self.L.orc_gen_range(_ptr(a), n, seed, stream, rng)
        return a

    def gen_perm(self, n, rng, seed, stream):
        a = np.empty(n, np.int32)
        self.L.orc_gen_perm(_ptr(a), n, rng, seed, stream)
        return a

    def zipf_cdf(self, n, s):
        T = np.empty(n, np.uint64)
        self.L.orc_zipf_cdf(_ptr(T), n, s)
        return T

    def gen_zipf(self, cnt, seed, stream, T, pseed=0, pstream=0):
        a = np.empty(cnt, np.int32)
        self.L.orc_gen_zipf(_ptr(a), cnt, seed, stream, _ptr(T), len(T), pseed, pstream)
        return a

    def fmix32(self, x):
        return self.L.orc_fmix32(x)

    # -------------------------------------------------------------- select
    def select(self, x, bound):
        out = np.empty(len(x), np.int32)
        k = self.L.orc_select(_ptr(x), len(x), bound, _ptr(out))
        return out[:k]

    def select_sum(self, pred, val, bound):
        c, s, s32 = U64(), C.c_int64(), I32()
        self.L.orc_select_sum(_ptr(pred), _ptr(val), len(pred), bound, C.byref(c), C.byref(s), C.byref(s32))
        return c.value, s.value, s32.value

    def select_sum_mt(self, pred, val, bound, threads):
        c, s = U64(), C.c_int64()
        self.L.orc_select_sum_mt(_ptr(pred), _ptr(val), len(pred), bound, C.byref(c), C.byref(s), threads)
        return c.value, s.value

    # -------------------------------------------------------------- join
    def table(self, keys, payload=None, log_buckets=None, threads=1):
        if log_buckets is None:
            log_buckets = max(4, int(np.ceil(np.log2(max(1, len(keys))))) + 1)
        t = ChainedTable()
        if threads > 1:
            rc = self.L.orc_ct_build_mt(C.byref(t), _ptr(keys), _ptr(payload), len(keys), log_buckets, threads)
        else:
            rc = self.L.orc_ct_build(C.byref(t), _ptr(keys), _ptr(payload), len(keys), log_buckets)
        assert rc == 0
        t._keep = (keys, payload)
        return t

    def free(self, t):
        self.L.orc_ct_free(C.byref(t))

    def probe(self, t, rk, rv=None, pred=None, bound=0, r_base=0, threads=1):
        r = JoinResult()
        if threads > 1:
            self.L.orc_ct_probe_mt(C.byref(t), _ptr(pred), bound, _ptr(rk), _ptr(rv), len(rk), C.byref(r), threads)
        else:
            self.L.orc_ct_probe(C.byref(t), _ptr(pred), bound, _ptr(rk), _ptr(rv), len(rk), r_base, C.byref(r))
        return {"count": r.count, "sum": r.sum, "digest_sum": r.digest_sum, "digest_xor": r.digest_xor,
                "count32": r.count32}

    def probe_groupby(self, t, rk, rv, pred, bound, ngroups, threads=1):
        cnt = np.zeros(ngroups, np.uint64)
        s = np.zeros(ngroups, np.int64)
        if threads > 1:
            rc = self.L.orc_ct_probe_groupby_mt(C.byref(t), _ptr(pred), bound, _ptr(rk), _ptr(rv), len(rk),
                                                _ptr(cnt), _ptr(s), ngroups, threads)
            assert rc == 0
        else:
            self.L.orc_ct_probe_groupby(C.byref(t), _ptr(pred), bound, _ptr(rk), _ptr(rv), len(rk), _ptr(cnt),
                                        _ptr(s), ngroups)
        return cnt, s

    def join_dense(self, sk, sw, kmax, rk, rv=None, pred=None, bound=0, threads=1):
        """(COUNT, SUM) of the join against distinct build keys in [0, kmax)
        by direct addressing; None when a build key is repeated or out of
        range."""
        c, s = U64(), C.c_int64()
        rc = self.L.orc_join_dense(_ptr(sk), _ptr(sw), len(sk), kmax, _ptr(pred), bound, _ptr(rk), _ptr(rv), len(rk),
                                   C.byref(c), C.byref(s), threads)
        return None if rc else (c.value, s.value)

    def groupby(self, keys, vals, ngroups):
        cnt = np.zeros(ngroups, np.uint64)
        s = np.zeros(ngroups, np.int64)
        self.L.orc_groupby(_ptr(keys), _ptr(vals), len(keys), _ptr(cnt), _ptr(s), ngroups)
        return cnt, s

    def selfjoin_count(self, x, bound, vmax):
        return self.L.orc_selfjoin_count(_ptr(x), len(x), bound, vmax)

    def split(self, cols, nparents, pred=None, bound=0):
        """Rows of each parent, in input order: a list of nparents lists of
        ncols int32 arrays."""
        cols = [np.ascontiguousarray(c, dtype=np.int32) for c in cols]
        n, nc = len(cols[0]), len(cols)
        outs = [np.zeros(max(n, 1), np.int32) for _ in range(nparents * nc)]
        cnt = np.zeros(nparents, np.uint64)
        cp = (VP * nc)(*[c.ctypes.data for c in cols])
        op = (VP * len(outs))(*[o.ctypes.data for o in outs])
        pr = np.ascontiguousarray(pred, dtype=np.int32) if pred is not None else None
        self.L.orc_split(_ptr(pr) if pr is not None else None, bound, cp, nc, n, nparents, op, _ptr(cnt))
        return [[outs[p * nc + c][: int(cnt[p])].copy() for c in range(nc)] for p in range(nparents)]


def col_check(x: np.ndarray) -> dict:
    """Same checksum as tests/golden/make_golden.py:col_sum."""
    from npref import mix64
    x = np.asarray(x)
    idx = np.arange(len(x), dtype=np.uint64)
    h = mix64((idx << np.uint64(32)) | x.astype(np.uint32).astype(np.uint64))
    with np.errstate(over="ignore"):
        hs = int(np.sum(h, dtype=np.uint64))
    return {"n": int(len(x)), "sum": int(x.astype(np.int64).sum()), "hash": hs}
