"""CPU tests: pin the oracle (oracle/mgv_oracle.c) against known answers,
library routines, closed forms, brute force and the golden fixtures written by
the independent numpy implementation (tests/golden/npref.py)."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import npref
from oracle_py import col_check

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ----------------------------------------------------------------- Philox

# Random123 known-answer vectors for philox4x32-10 (kat_vectors).
R123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", R123_KAT)
def test_philox_random123_kat(orc, ctr, key, out):
    assert orc.philox(ctr, key) == out
    got = npref.philox(np.array([ctr], dtype=np.uint32), key[0], key[1])[0]
    assert [int(v) for v in got] == out


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available to build the rocRAND cross-check")
def test_philox_matches_rocrand_engine(orc):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "rp")
        subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1",
                        os.path.join(ROOT, "tests", "rocrand_philox.cpp"), "-o", exe],
                       check=True, capture_output=True, cwd=d)
        for seed, stream in [(orc.seed(1, 1), 0), (orc.seed(2, 3), 7), (0xDEADBEEFCAFEF00D, 123456)]:
            blks = [0, 1, 2, 1000, 2**32 + 5]
            out = subprocess.run([exe, hex(seed), str(stream)] + [str(b) for b in blks], check=True,
                                 capture_output=True, text=True).stdout.split("\n")
            for b, line in zip(blks, out):
                want = [int(v) for v in line.split()]
                got = [orc.word(seed, stream, 4 * b + j) for j in range(4)]
                assert got == want, (seed, stream, b)


def test_words_golden(orc, golden):
    for w in golden["words"]:
        assert orc.word(w["seed"], w["stream"], w["i"]) == w["word"]


# ----------------------------------------------------------------- fmix32

def _murmur3_x86_32_body(key: int, seed: int) -> int:
    """MurmurHash3_x86_32 of a 4-byte key up to (excluding) the fmix32 step."""
    M = 0xFFFFFFFF
    rotl = lambda x, r: ((x << r) | (x >> (32 - r))) & M
    k = (key & M) * 0xCC9E2D51 & M
    k = rotl(k, 15)
    k = k * 0x1B873593 & M
    h = (seed & M) ^ k
    h = rotl(h, 13)
    h = (h * 5 + 0xE6546B64) & M
    return h ^ 4


def test_fmix32_matches_sklearn_murmurhash3(orc):
    from sklearn.utils import murmurhash3_32
    rng = np.random.default_rng(7)
    for key, seed in zip(rng.integers(-2**31, 2**31, 200), rng.integers(0, 2**31, 200)):
        want = murmurhash3_32(int(key), seed=int(seed), positive=True)
        assert orc.fmix32(_murmur3_x86_32_body(int(key), int(seed))) == want


def test_fmix32_golden_and_bijective(orc, golden):
    for e in golden["fmix32"]:
        assert orc.fmix32(e["x"]) == e["h"]
    xs = np.random.default_rng(1).integers(0, 2**32, 200000, dtype=np.uint64)
    xs = np.unique(xs)
    hs = npref.fmix32(xs)
    assert len(np.unique(hs)) == len(xs)
    assert orc.fmix32(0) == 0


# ----------------------------------------------------------------- Feistel

def test_perm_is_bijection_and_golden(orc, golden):
    for e in golden["perm"]:
        p = orc.gen_perm(e["range"], e["range"], e["seed"], e["stream"])
        assert sorted(p.tolist()) == list(range(e["range"]))
        assert p[:16].tolist() == e["head"]
        assert col_check(p) == e["check"]
    # power-of-four ranges are a single Feistel pass; others cycle-walk
    p = orc.gen_perm(1 << 16, 1 << 16, orc.seed(2, 0), 3)
    assert len(np.unique(p)) == 1 << 16
    assert not np.array_equal(p, np.arange(1 << 16))


# ----------------------------------------------------------------- Zipf

def test_zipf_cdf_closed_form_s1(orc):
    n = 64
    T = orc.zipf_cdf(n, 1.0)
    Hn = sum(Fraction(1, k) for k in range(1, n + 1))
    acc = Fraction(0)
    for r in range(n):
        acc += Fraction(1, r + 1)
        exact = (acc / Hn * 2**32).__floor__()
        if r == n - 1:
            exact = 2**32
        assert abs(int(T[r]) - exact) <= 1, r


def test_zipf_golden_and_sampling(orc, golden):
    for e in golden["zipf"]:
        T = orc.zipf_cdf(e["n"], e["s"])
        assert T[:8].tolist() == e["head"]
        assert int(T.sum()) == e["sum"]
        assert int(T[e["n"] // 2]) == e["mid"]
    T = orc.zipf_cdf(1024, 1.0)
    x = orc.gen_zipf(200000, orc.seed(3, 1), 0, T)
    H = sum(1.0 / k for k in range(1, 1025))
    f0 = np.mean(x == 0)
    assert abs(f0 - 1.0 / H) < 0.01
    assert x.min() >= 0 and x.max() < 1024


# ----------------------------------------------------------------- datagen vs golden

def test_datagen_matches_numpy(orc, golden):
    c1 = golden["c1_small"]
    n = c1["n"]
    a = orc.gen_bits(n, orc.seed(1, 0), 0, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 0, 31)
    assert col_check(a) == c1["a"] and col_check(b) == c1["b"]
    assert np.array_equal(a, npref.gen_bits(n, npref.seed(1, 0), 0, 20))
    r = orc.gen_range(1000, orc.seed(0, 0), 0, 10000)
    assert np.array_equal(r, npref.gen_range(1000, npref.seed(0, 0), 0, 10000))
    assert r.min() >= 0 and r.max() < 10000


# ----------------------------------------------------------------- select / sum

def test_select_sum_golden(orc, golden):
    c1 = golden["c1_small"]
    n = c1["n"]
    a = orc.gen_bits(n, orc.seed(1, 0), 0, 20)
    b = orc.gen_bits(n, orc.seed(1, 1), 0, 31)
    cnt, s64, s32 = orc.select_sum(b, a, c1["bound"])
    assert (cnt, s64, s32) == (c1["count"], c1["sum64"], c1["sum32"])
    assert np.int64(s64).astype(np.int32) == s32          # reference int32 wrap semantics
    # a planted off-by-one in the bound must change the answer
    assert orc.select_sum(b, a, c1["bound"] + 1) != (cnt, s64, s32) or not np.any(b == c1["bound"] + 1)
    assert orc.select_sum_mt(b, a, c1["bound"], 4) == (cnt, s64)


def test_reference_single_column_sum_and_select(orc, golden):
    e = golden["ref_sum_select"]
    x = orc.gen_range(e["n"], orc.seed(0, 0), 0, 10000) + 1
    assert col_check(x) == e["x"]
    cnt, s64, s32 = orc.select_sum(x, x, e["bound"])
    assert (cnt, s64, s32) == (e["count"], e["sum64"], e["sum32"])
    sel = orc.select(x, e["bound"])
    assert np.array_equal(sel, x[x <= e["bound"]])        # stable_select_cpu keeps order
    assert orc.select(x, -1).size == 0 and orc.select(x, 2**31 - 1).size == len(x)


def test_selfjoin_hist_squares(orc, golden):
    e = golden["ref_selfjoin"]
    x = orc.gen_range(e["n"], orc.seed(0, 0), 0, 10000) + 1
    assert orc.selfjoin_count(x, e["bound"], 10000) == e["count"]
    f = orc.select(x, e["bound"])
    t = orc.table(f)
    assert orc.probe(t, f)["count"] == e["count"]           # chained join of the column with itself
    orc.free(t)


# ----------------------------------------------------------------- join

def _c2_small(orc, g):
    sk = orc.gen_perm(g["ns"], g["ns"], orc.seed(2, 0), 0)
    sw = orc.gen_bits(g["ns"], orc.seed(2, 1), 0, 16)
    rk = orc.gen_bits(g["nr"], orc.seed(2, 2), 0, 13)
    rv = orc.gen_bits(g["nr"], orc.seed(2, 3), 0, 16)
    return sk, sw, rk, rv


def test_join_golden_and_closed_form(orc, golden):
    g = golden["c2_small"]
    sk, sw, rk, rv = _c2_small(orc, g)
    for name, col in (("sk", sk), ("sw", sw), ("rk", rk), ("rv", rv)):
        assert col_check(col) == g[name]
    t = orc.table(sk, sw)
    r = orc.probe(t, rk, rv)
    assert r["count"] == g["count"] == int((rk < g["ns"]).sum())
    assert r["sum"] == g["sum"]
    assert r["count32"] == np.int64(r["count"]).astype(np.int32)
    assert orc.probe(t, rk, rv, threads=3)["sum"] == g["sum"]
    orc.free(t)
    t = orc.table(sk, np.arange(g["ns"], dtype=np.int32))
    d = orc.probe(t, rk)
    assert (d["digest_sum"], d["digest_xor"]) == (g["digest_sum"], g["digest_xor"])
    orc.free(t)


def test_join_bucket_count_does_not_change_result(orc, golden):
    g = golden["c2_small"]
    sk, sw, rk, rv = _c2_small(orc, g)
    want = None
    for lb in (1, 5, 12, 16):
        t = orc.table(sk, sw, log_buckets=lb)
        r = orc.probe(t, rk, rv)
        orc.free(t)
        want = want or r
        assert r == want


def test_join_bruteforce_tiny(orc):
    rng = np.random.default_rng(3)
    for _ in range(5):
        ns, nr = int(rng.integers(0, 60)), int(rng.integers(0, 90))
        sk = rng.integers(-5, 20, ns).astype(np.int32)
        sw = rng.integers(-1000, 1000, ns).astype(np.int32)
        rk = rng.integers(-5, 25, nr).astype(np.int32)
        rv = rng.integers(-1000, 1000, nr).astype(np.int32)
        pred = rng.integers(0, 10, nr).astype(np.int32)
        cnt = tot = 0
        for r in range(nr):
            if pred[r] > 4:
                continue
            for s in range(ns):
                if sk[s] == rk[r]:
                    cnt += 1
                    tot += int(rv[r]) + int(sw[s])
        t = orc.table(sk, sw)
        got = orc.probe(t, rk, rv, pred=pred, bound=4)
        orc.free(t)
        assert (got["count"], got["sum"]) == (cnt, tot)


def test_duplicate_keys_golden(orc, golden):
    g = golden["dup_join"]
    sk = orc.gen_range(g["ns"], orc.seed(2, 10), 0, 64)
    sw = orc.gen_bits(g["ns"], orc.seed(2, 11), 0, 16)
    rk = orc.gen_range(g["nr"], orc.seed(2, 12), 0, 80)
    rv = orc.gen_bits(g["nr"], orc.seed(2, 13), 0, 16)
    assert col_check(sk) == g["sk"] and col_check(rk) == g["rk"]
    t = orc.table(sk, sw)
    r = orc.probe(t, rk, rv)
    orc.free(t)
    assert (r["count"], r["sum"]) == (g["count"], g["sum"])
    hs = np.bincount(sk, minlength=80)
    assert r["count"] == int(sum(hs[k] for k in rk))
    t = orc.table(sk, np.arange(g["ns"], dtype=np.int32))
    d = orc.probe(t, rk)
    orc.free(t)
    assert (d["digest_sum"], d["digest_xor"]) == (g["digest_sum"], g["digest_xor"])


# ----------------------------------------------------------------- group-by

def test_select_join_groupby_golden(orc, golden):
    g = golden["c3_small"]
    T = orc.zipf_cdf(g["nranks"], 1.0)
    c = orc.gen_bits(g["nr"], orc.seed(3, 0), 0, 31)
    rk = orc.gen_zipf(g["nr"], orc.seed(3, 1), 0, T, orc.seed(3, 5), 0)
    rv = orc.gen_bits(g["nr"], orc.seed(3, 2), 0, 16)
    sk = orc.gen_perm(g["nranks"], g["nranks"], orc.seed(3, 3), 0)
    sg = orc.gen_bits(g["nranks"], orc.seed(3, 4), 0, 6)
    for name, col in (("c", c), ("rk", rk), ("rv", rv), ("sk", sk), ("sg", sg)):
        assert col_check(col) == g[name], name
    tg = orc.table(sk, sg)
    tn = orc.table(sk)
    for case in g["cases"]:
        cnt, s = orc.probe_groupby(tg, rk, rv, c, case["bound"], g["ngroups"])
        assert cnt.tolist() == case["gcnt"] and s.tolist() == case["gsum"]
        r = orc.probe(tn, rk, rv, pred=c, bound=case["bound"])
        assert (r["count"], r["sum"]) == (case["count"], case["sum"])
        assert int(cnt.sum()) == r["count"] and int(s.sum()) == r["sum"]
    orc.free(tg)
    orc.free(tn)


def test_groupby_golden(orc, golden):
    g = golden["groupby_small"]
    T = orc.zipf_cdf(g["ngroups"], 0.8)
    k = orc.gen_zipf(g["n"], orc.seed(5, 0), 0, T)
    v = orc.gen_bits(g["n"], orc.seed(5, 1), 0, 16)
    assert col_check(k) == g["k"] and col_check(v) == g["v"]
    cnt, s = orc.groupby(k, v, g["ngroups"])
    assert col_check(cnt.astype(np.int64)) == g["gcnt_hash"]
    assert col_check(s) == g["gsum_hash"]
    assert int(s.sum()) == g["total_sum"] == int(v.astype(np.int64).sum())


# ----------------------------------------------------------------- split

def _rows(cols):
    """Rows of a column list as a sorted list of tuples (a multiset)."""
    return sorted(zip(*[c.tolist() for c in cols])) if len(cols[0]) else []


@pytest.mark.parametrize("nparents", [1, 2, 3, 8])
@pytest.mark.parametrize("ncols,filtered", [(1, False), (2, True), (4, False), (3, True)])
def test_split_partitions_the_kept_rows_by_parity(orc, nparents, ncols, filtered):
    n = 5000
    rng = np.random.default_rng(11 + ncols)
    cols = [rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32) for _ in range(ncols)]
    cols[0][:7] = [0, -1, 1, -2, 2**31 - 1, -2**31, 3]     # signs and extremes of the parity test
    pred = rng.integers(0, 100, n).astype(np.int32) if filtered else None
    bound = 49
    parts = orc.split(cols, nparents, pred=pred, bound=bound)
    keep = np.ones(n, bool) if pred is None else pred <= bound
    par = np.array([int(x) % 2 for x in cols[0]])           # Python modulo: independent of the C & 1
    for p in range(nparents):
        sel = keep & (par == p)
        if p >= 2:
            assert len(parts[p][0]) == 0                     # x & 1 never names parent 2 or more
            continue
        # parent p: exactly the kept rows of that parity, in input order
        for c in range(ncols):
            assert np.array_equal(parts[p][c], cols[c][sel])
    got = sorted(_rows(parts[0]) + (_rows(parts[1]) if nparents > 1 else []))
    want = _rows([c[keep & ((par == 0) | (nparents > 1))] for c in cols])
    assert got == want                                       # nothing lost or duplicated


def test_split_of_nothing(orc):
    parts = orc.split([np.zeros(0, np.int32)], 2)
    assert [len(p[0]) for p in parts] == [0, 0]


# ----------------------------------------------------------------- mix64

# SplitMix64 (Steele, Lea, Flood, OOPSLA'14; reference C code by S. Vigna):
# state += 0x9e3779b97f4a7c15, output = mix64(state).  Published output
# sequences of the reference generator for two seeds.
SPLITMIX64_KAT = [
    (1234567, [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
               16408922859458223821]),
    (1477776061723855037, [1985237415132408290, 2979275885539914483, 13511426838097143398]),
]


@pytest.mark.parametrize("seed,outs", SPLITMIX64_KAT)
def test_mix64_is_the_splitmix64_finaliser(orc, seed, outs):
    s = seed
    for want in outs:
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        assert orc.L.orc_mix64(s) == want


# ----------------------------------------------------------------- Zipf, s != 1

@pytest.mark.parametrize("n,s", [(1 << 16, 0.8), (4096, 1.3), (1000, 0.5)])
def test_zipf_cdf_weights_and_normaliser(orc, n, s):
    """Threshold gaps are the Zipf weights: (T[k] - T[k-1]) / T[0] = (k+1)^-s,
    and T[0] / 2^32 = 1 / H(n, s) with the generalized harmonic number from
    the Euler-Maclaurin expansion around Riemann's zeta (scipy), not from the
    sum the oracle adds up."""
    from scipy.special import zeta
    T = orc.zipf_cdf(n, s).astype(np.float64)
    w = np.diff(np.concatenate([[0.0], T]))
    k = np.array([1, 2, 3, 7, 100, n // 2, n - 2])
    # each threshold is floored, so a gap is within 2 of its exact value
    assert np.all(np.abs(w[k] - w[0] * (k + 1.0) ** -s) <= 2.5)
    assert T[-1] == 2.0**32
    assert np.all(np.diff(T) >= 0)
    H = zeta(s) + n ** (1 - s) / (1 - s) + n**-s / 2 - s * n ** (-s - 1) / 12
    assert abs(T[0] / 2**32 - 1.0 / H) < 1e-7


# ----------------------------------------------------------------- direct-address join, threaded oracle

def test_join_dense_equals_chained_join_and_closed_form(orc, golden):
    g = golden["c2_small"]
    sk, sw, rk, rv = _c2_small(orc, g)
    t = orc.table(sk, sw)
    c = orc.gen_bits(len(rk), orc.seed(3, 0), 0, 31)
    for pred, bound in ((None, 0), (c, (1 << 30) - 1), (c, -1)):
        want = orc.probe(t, rk, rv, pred=pred, bound=bound)
        for threads in (1, 3):
            got = orc.join_dense(sk, sw, g["ns"], rk, rv, pred=pred, bound=bound, threads=threads)
            assert got == (want["count"], want["sum"])
    orc.free(t)
    assert orc.join_dense(sk, sw, g["ns"], rk, rv)[0] == int((rk < g["ns"]).sum()) == g["count"]
    assert orc.join_dense(sk, None, g["ns"], rk, None) == (g["count"], 0)


def test_join_dense_rejects_repeated_or_out_of_range_keys(orc):
    rk = np.arange(10, dtype=np.int32)
    assert orc.join_dense(np.array([1, 2, 1], np.int32), None, 10, rk) is None
    assert orc.join_dense(np.array([1, 12], np.int32), None, 10, rk) is None
    assert orc.join_dense(np.array([-1], np.int32), None, 10, rk) is None
    assert orc.join_dense(np.array([3, 9], np.int32), np.array([5, -7], np.int32), 10, rk, rk) == (2, 3 + 5 + 9 - 7)


def test_threaded_build_and_groupby_equal_serial(orc):
    ns, nr, G = 50000, 200000, 97
    sk = orc.gen_range(ns, orc.seed(2, 10), 4, 20000)       # repeated keys
    sg = orc.gen_range(ns, orc.seed(2, 11), 4, G)
    rk = orc.gen_range(nr, orc.seed(2, 12), 4, 25000)
    rv = orc.gen_bits(nr, orc.seed(2, 13), 4, 16)
    c = orc.gen_bits(nr, orc.seed(3, 0), 4, 31)
    t1 = orc.table(sk, sg, log_buckets=14)
    t8 = orc.table(sk, sg, log_buckets=14, threads=8)
    # the same rows hang off every bucket
    for b in range(0, 1 << 14, 997):
        rows = []
        for t in (t1, t8):
            heads = np.ctypeslib.as_array((np.ctypeslib.ctypes.c_int32 * (1 << 14)).from_address(t.heads))
            nxt = np.ctypeslib.as_array((np.ctypeslib.ctypes.c_int32 * ns).from_address(t.next))
            cur, chain = int(heads[b]), []
            while cur >= 0:
                chain.append(cur)
                cur = int(nxt[cur])
            rows.append(sorted(chain))
        assert rows[0] == rows[1]
    w1 = orc.probe(t1, rk, rv, pred=c, bound=(1 << 30) - 1)
    assert orc.probe(t8, rk, rv, pred=c, bound=(1 << 30) - 1, threads=5) == w1
    g1 = orc.probe_groupby(t1, rk, rv, c, (1 << 30) - 1, G)
    g8 = orc.probe_groupby(t8, rk, rv, c, (1 << 30) - 1, G, threads=6)
    assert np.array_equal(g1[0], g8[0]) and np.array_equal(g1[1], g8[1])
    assert int(g1[0].sum()) == w1["count"]
    orc.free(t1)
    orc.free(t8)


# ----------------------------------------------------------------- fmix32 inverse

def test_fmix32_inverse_constants_undo_the_oracle_fmix32(orc):
    """The library recovers a key from its hash (mgv_wave.hpp fmix32_inv) with
    the multiplicative inverses 0xa5cb9243 = 0x85ebca6b^-1 and 0x7ed1b41d =
    0xc2b2ae35^-1 mod 2^32 and the xor-shift inverses; composed with the
    oracle's fmix32 (pinned by sklearn's murmurhash3 above) the restatement
    must give back every input."""
    assert (0xa5cb9243 * 0x85ebca6b) % 2**32 == 1 and (0x7ed1b41d * 0xc2b2ae35) % 2**32 == 1
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.integers(0, 2**32, 1 << 16, dtype=np.uint64),
                        np.array([0, 1, 2**31, 2**32 - 1, 0x85ebca6b], dtype=np.uint64)]).astype(np.uint32)
    h = np.array([orc.L.orc_fmix32(int(v)) for v in x], dtype=np.uint64)
    M = np.uint64(0xFFFFFFFF)
    y = h ^ (h >> np.uint64(16))
    y = (y * np.uint64(0x7ed1b41d)) & M
    y = y ^ (y >> np.uint64(13)) ^ (y >> np.uint64(26))
    y = (y * np.uint64(0xa5cb9243)) & M
    y = y ^ (y >> np.uint64(16))
    assert np.array_equal(y.astype(np.uint32), x)
