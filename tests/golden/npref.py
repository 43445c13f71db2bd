"""Independent numpy implementation of the synthetic-data generator and of the
query semantics, used to write and check the golden fixtures.

It shares no code with the C oracle (oracle/mgv_oracle.c) or the HIP library:
join results come from direct addressing / np.unique instead of a hash table,
so a fixture that both the oracle and this module agree on pins the oracle.
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)


def seed(config: int, column: int) -> int:
    return 0x4D47560000000000 | (config << 8) | column


def philox(ctr: np.ndarray, key0: int, key1: int) -> np.ndarray:
    """Philox4x32-10 over an (n, 4) uint32 counter array -> (n, 4) uint32."""
    c = ctr.astype(np.uint64)
    k0, k1 = key0 & 0xFFFFFFFF, key1 & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[:, 0]
        p1 = M1 * c[:, 2]
        n0 = (p1 >> np.uint64(32)) ^ c[:, 1] ^ np.uint64(k0)
        n1 = p1 & MASK32
        n2 = (p0 >> np.uint64(32)) ^ c[:, 3] ^ np.uint64(k1)
        n3 = p0 & MASK32
        c = np.stack([n0, n1, n2, n3], axis=1)
    return c.astype(np.uint32)


def words(n: int, seed_: int, stream: int = 0) -> np.ndarray:
    nblk = (n + 3) // 4
    blk = np.arange(nblk, dtype=np.uint64)
    ctr = np.zeros((nblk, 4), dtype=np.uint64)
    ctr[:, 0] = blk & MASK32
    ctr[:, 1] = blk >> np.uint64(32)
    ctr[:, 2] = stream
    out = philox(ctr, seed_ & 0xFFFFFFFF, seed_ >> 32)
    return out.reshape(-1)[:n]


def gen_bits(n: int, seed_: int, stream: int, bits: int) -> np.ndarray:
    w = words(n, seed_, stream).astype(np.uint64)
    if bits == 0:
        return np.zeros(n, dtype=np.int32)
    return (w >> np.uint64(32 - bits)).astype(np.uint32).view(np.int32)


def gen_range(n: int, seed_: int, stream: int, rng: int) -> np.ndarray:
    w = words(n, seed_, stream).astype(np.uint64)
    return ((w * np.uint64(rng)) >> np.uint64(32)).astype(np.uint32).view(np.int32)


def fmix32(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64) & MASK32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x85EBCA6B)) & MASK32
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0xC2B2AE35)) & MASK32
    x ^= x >> np.uint64(16)
    return x


def mix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def _round_key(seed_: int, stream: int, r: int) -> int:
    inner = ((seed_ >> 32) + stream * 0x85EBCA6B + r * 0x9E3779B9) & 0xFFFFFFFF
    return int(fmix32((seed_ & 0xFFFFFFFF) ^ int(fmix32(inner))))


def feistel(x: np.ndarray, nbits: int, seed_: int, stream: int) -> np.ndarray:
    h = nbits // 2
    m = np.uint64((1 << h) - 1)
    x = np.asarray(x, dtype=np.uint64)
    L, R = x >> np.uint64(h), x & m
    for r in range(4):
        F = fmix32(R ^ np.uint64(_round_key(seed_, stream, r))) & m
        L, R = R, L ^ F
    return (L << np.uint64(h)) | R


def perm_width(rng: int) -> int:
    nb = 2
    while nb < 32 and (1 << nb) < rng:
        nb += 2
    return nb


def perm(x: np.ndarray, rng: int, seed_: int, stream: int) -> np.ndarray:
    nb = perm_width(rng)
    y = feistel(np.asarray(x, dtype=np.uint64), nb, seed_, stream)
    todo = y >= rng
    while todo.any():
        y[todo] = feistel(y[todo], nb, seed_, stream)
        todo = y >= rng
    return y.astype(np.uint32).view(np.int32)


def zipf_cdf(n: int, s: float) -> np.ndarray:
    r = np.arange(1, n + 1, dtype=np.float64)
    w = 1.0 / r if s == 1.0 else 1.0 / np.power(r, s)
    H = 0.0
    for v in w:  # sequential, ascending, like the C code
        H += v
    acc = np.cumsum(w)  # cumsum is sequential in numpy
    f = acc / H * 4294967296.0
    T = np.where(f >= 4294967296.0, 4294967296, np.floor(f)).astype(np.uint64)
    T[-1] = 4294967296
    return T


def gen_zipf(cnt: int, seed_: int, stream: int, T: np.ndarray, pseed: int = 0, pstream: int = 0) -> np.ndarray:
    u = words(cnt, seed_, stream).astype(np.uint64)
    rank = np.searchsorted(T, u, side="right")  # smallest r with u < T[r]
    if pseed:
        return perm(rank, len(T), pseed, pstream)
    return rank.astype(np.int32)


# ------------------------------------------------------------------ semantics


def select_sum(pred, val, bound):
    sel = pred <= bound
    cnt = int(sel.sum())
    s64 = int(val[sel].astype(np.int64).sum())
    s32 = int(np.int32(np.uint32(s64 & 0xFFFFFFFF).view(np.int32)))
    return cnt, s64, s32


def join_pairs(sk, rk):
    """All (r, s) index pairs with rk[r] == sk[s], via sorting (no hashing)."""
    order = np.argsort(sk, kind="stable")
    sks = sk[order]
    lo = np.searchsorted(sks, rk, side="left")
    hi = np.searchsorted(sks, rk, side="right")
    cnts = hi - lo
    r_idx = np.repeat(np.arange(len(rk)), cnts)
    starts = np.repeat(lo, cnts)
    within = np.arange(len(r_idx)) - np.repeat(np.cumsum(cnts) - cnts, cnts)
    s_idx = order[starts + within]
    return r_idx.astype(np.int64), s_idx.astype(np.int64)


def join_agg(sk, sw, rk, rv, pred=None, bound=None, r_base=0):
    keep = np.ones(len(rk), dtype=bool) if pred is None else (pred <= bound)
    r_idx, s_idx = join_pairs(sk, rk)
    m = keep[r_idx]
    r_idx, s_idx = r_idx[m], s_idx[m]
    cnt = len(r_idx)
    sw64 = np.zeros(len(sk), np.int64) if sw is None else sw.astype(np.int64)
    rv64 = np.zeros(len(rk), np.int64) if rv is None else rv.astype(np.int64)
    total = int(rv64[r_idx].sum() + sw64[s_idx].sum())
    d = mix64((r_idx.astype(np.uint64) + np.uint64(r_base)) << np.uint64(32) | s_idx.astype(np.uint64))
    with np.errstate(over="ignore"):
        dsum = int(np.sum(d, dtype=np.uint64))
    dxor = int(np.bitwise_xor.reduce(d)) if len(d) else 0
    return {"count": cnt, "sum": total, "digest_sum": dsum, "digest_xor": dxor}


def groupby(keys, vals, ngroups):
    k = keys.astype(np.int64)
    ok = (k >= 0) & (k < ngroups)
    cnt = np.bincount(k[ok], minlength=ngroups).astype(np.uint64)
    s = np.bincount(k[ok], weights=None, minlength=ngroups)
    sums = np.zeros(ngroups, dtype=np.int64)
    np.add.at(sums, k[ok], vals[ok].astype(np.int64))
    return cnt, sums


def join_groupby(sk, sg, rk, rv, pred, bound, ngroups):
    keep = pred <= bound
    r_idx, s_idx = join_pairs(sk, rk)
    m = keep[r_idx]
    return groupby(sg[s_idx[m]], rv[r_idx[m]], ngroups)
