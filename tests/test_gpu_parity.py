"""GPU parity tests: libmgv.so (through its C ABI) against the CPU oracle and
the golden fixtures, on the same seeded inputs.  Integer results must be
bit-exact."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv
from oracle_py import col_check

pytestmark = pytest.mark.gpu

INT_MIN, INT_MAX = -2**31, 2**31 - 1


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def gen(gpu, kind, n, *args):
    buf = mgv.DeviceBuffer(4 * max(n, 1))
    gpu.call({"bits": "mgv_gen_bits", "range": "mgv_gen_range", "perm": "mgv_gen_perm",
              "iota": "mgv_gen_iota"}[kind], buf.vp, n, *args, None)
    return buf


def down(buf, n):
    gpu_sync()
    return buf.download(np.int32, n)


def gpu_sync():
    mgv.sync()


# ----------------------------------------------------------------- datagen

@pytest.mark.parametrize("n", [1, 5, 4096, 16461])
def test_datagen_matches_oracle(gpu, orc, n):
    s = orc.seed(1, 1)
    assert np.array_equal(down(gen(gpu, "bits", n, s, 3, 31), n), orc.gen_bits(n, s, 3, 31))
    assert np.array_equal(down(gen(gpu, "bits", n, s, 0, 0), n), np.zeros(n, np.int32))
    assert np.array_equal(down(gen(gpu, "range", n, s, 1, 10000), n), orc.gen_range(n, s, 1, 10000))
    rng = max(n, 37)
    assert np.array_equal(down(gen(gpu, "perm", n, rng, s, 2), n), orc.gen_perm(n, rng, s, 2))
    assert np.array_equal(down(gen(gpu, "iota", n, -5), n), np.arange(-5, n - 5, dtype=np.int32))


def test_zipf_datagen_matches_oracle(gpu, orc):
    for nranks, sexp, pseed in [(1024, 1.0, orc.seed(3, 5)), (4096, 0.8, 0)]:
        T = C.c_void_p()
        gpu.call("mgv_zipf_table_create", C.byref(T), nranks, sexp)
        Tb = mgv.DeviceBuffer.__new__(mgv.DeviceBuffer)
        Tb.ptr, Tb.nbytes = T.value, 8 * nranks
        assert np.array_equal(Tb.download(np.uint64, nranks), orc.zipf_cdf(nranks, sexp))
        n = 10001
        out = mgv.DeviceBuffer(4 * n)
        gpu.call("mgv_gen_zipf", out.vp, n, orc.seed(3, 1), 0, Tb.vp, nranks, pseed, 0, None)
        want = orc.gen_zipf(n, orc.seed(3, 1), 0, orc.zipf_cdf(nranks, sexp), pseed, 0)
        assert np.array_equal(down(out, n), want)
        Tb.close()


# ----------------------------------------------------------------- select + sum

def run_select_sum(gpu, pred, val, n, bound, launch):
    res = mgv.ResultBuffer()
    gpu.call("mgv_select_sum", pred.vp, val.vp, n, bound, res.vp, launch, None)
    gpu_sync()
    return res.read()


@pytest.mark.parametrize("block", [64, 256, 1024])
@pytest.mark.parametrize("compact", [0, 1])
def test_select_sum_c1_small(gpu, orc, golden, block, compact):
    g = golden["c1_small"]
    n = g["n"]
    a = gen(gpu, "bits", n, orc.seed(1, 0), 0, 20)
    b = gen(gpu, "bits", n, orc.seed(1, 1), 0, 31)
    r = run_select_sum(gpu, b, a, n, g["bound"], mgv.launch(0, block, compact))
    assert (r.count, r.sum) == (g["count"], g["sum64"])
    assert np.int64(r.sum).astype(np.int32) == g["sum32"]     # reference int32 semantics


def test_select_sum_reference_single_column(gpu, orc, golden):
    e = golden["ref_sum_select"]
    x = orc.gen_range(e["n"], orc.seed(0, 0), 0, 10000) + 1
    d = dev(x)
    r = run_select_sum(gpu, d, d, e["n"], e["bound"], mgv.launch(128, 1024, 1))
    assert (r.count, r.sum) == (e["count"], e["sum64"])


@pytest.mark.parametrize("n", [0, 1, 3, 7, 255, 257, 1023, 100003])
def test_select_sum_edges(gpu, orc, n):
    rng = np.random.default_rng(n)
    pred = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64, endpoint=True).astype(np.int32)
    val = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64, endpoint=True).astype(np.int32)
    if n:
        pred[0], pred[-1] = INT_MIN, INT_MAX
    dp, dv = dev(pred), dev(val)
    for bound in (INT_MIN, -1, 0, 12345, INT_MAX):
        for compact in (0, 1):
            r = run_select_sum(gpu, dp, dv, n, bound, mgv.launch(0, 256, compact))
            c, s, _ = orc.select_sum(pred, val, bound)
            assert (r.count, r.sum) == (c, s), (n, bound, compact)


def test_plain_per_vector_equals_resident(gpu, orc, golden):
    """One launch per 4096-row vector with the reference's 128 x 1024 launch
    shape accumulates to the same result as one resident launch."""
    g = golden["c1_small"]
    n = g["n"]
    a = gen(gpu, "bits", n, orc.seed(1, 0), 0, 20)
    b = gen(gpu, "bits", n, orc.seed(1, 1), 0, 31)
    res = mgv.ResultBuffer()
    V = 4096
    for off in range(0, n, V):
        m = min(V, n - off)
        gpu.call("mgv_select_sum", b.at(4 * off), a.at(4 * off), m, g["bound"], res.vp, mgv.launch(128, 1024, 1),
                 None)
    gpu_sync()
    r = res.read()
    assert (r.count, r.sum) == (g["count"], g["sum64"])


def test_select_materialize_multiset(gpu, orc):
    n = 50000
    x = orc.gen_range(n, orc.seed(0, 0), 0, 10000) + 1
    d = dev(x)
    out = mgv.DeviceBuffer(4 * n)
    res = mgv.ResultBuffer()
    gpu.call("mgv_select_materialize", d.vp, d.vp, n, 5000, out.vp, n, res.vp, mgv.launch(0, 256), None)
    gpu_sync()
    r = res.read()
    want = orc.select(x, 5000)
    assert r.count == len(want) and r.overflow == 0
    assert np.array_equal(np.sort(out.download(np.int32, r.count)), np.sort(want))
    # capacity smaller than the output: written prefix is a sub-multiset, the rest counted
    res.clear()
    cap = 1000
    gpu.call("mgv_select_materialize", d.vp, d.vp, n, 5000, out.vp, cap, res.vp, mgv.launch(0, 256), None)
    gpu_sync()
    r = res.read()
    assert r.count == len(want) and r.overflow == len(want) - cap
    got = out.download(np.int32, cap)
    assert np.all(got <= 5000) and np.all(got >= 1)


# ----------------------------------------------------------------- hash join

def build(gpu, keys, payload, n, lf=0.5, mode=mgv.MGV_PAYLOAD_COLUMN, pred=None, bound=0, block=256):
    ht = mgv.HashTable(n, lf)
    res = mgv.ResultBuffer()
    gpu.call("mgv_hashjoin_build", ht.buf.vp, ht.nbuckets, keys.vp, payload.vp if payload is not None else None, mode,
             pred.vp if pred is not None else None, bound, n, 0, res.vp, mgv.launch(0, block), None)
    gpu_sync()
    return ht, res.read()


def probe(gpu, ht, keys, vals, n, pred=None, bound=0, flags=0, launch=None):
    res = mgv.ResultBuffer()
    gpu.call("mgv_hashjoin_probe", ht.buf.vp, ht.nbuckets, pred.vp if pred is not None else None, bound, keys.vp,
             vals.vp if vals is not None else None, n, 0, flags, res.vp, launch or mgv.launch(0, 256), None)
    gpu_sync()
    return res.read()


def c2_small_cols(gpu, orc, g):
    sk = gen(gpu, "perm", g["ns"], g["ns"], orc.seed(2, 0), 0)
    sw = gen(gpu, "bits", g["ns"], orc.seed(2, 1), 0, 16)
    rk = gen(gpu, "bits", g["nr"], orc.seed(2, 2), 0, 13)
    rv = gen(gpu, "bits", g["nr"], orc.seed(2, 3), 0, 16)
    return sk, sw, rk, rv


@pytest.mark.parametrize("lf", [0.25, 0.5, 0.75, 0.9])
def test_join_c2_small(gpu, orc, golden, lf):
    g = golden["c2_small"]
    sk, sw, rk, rv = c2_small_cols(gpu, orc, g)
    ht, br = build(gpu, sk, sw, g["ns"], lf)
    assert br.errors == 0
    slots = ht.slots()
    assert int((slots != np.uint64(2**64 - 1)).sum()) == g["ns"]
    r = probe(gpu, ht, rk, rv, g["nr"])
    assert (r.count, r.sum) == (g["count"], g["sum"])
    # digest of the match multiset with payload = S row id
    ht2, _ = build(gpu, sk, None, g["ns"], lf, mgv.MGV_PAYLOAD_ROWID)
    d = probe(gpu, ht2, rk, None, g["nr"], flags=mgv.MGV_PROBE_DIGEST)
    assert (d.count, d.digest_sum, d.digest_xor) == (g["count"], g["digest_sum"], g["digest_xor"])


def test_join_duplicate_keys(gpu, orc, golden):
    g = golden["dup_join"]
    sk = dev(orc.gen_range(g["ns"], orc.seed(2, 10), 0, 64))
    sw = dev(orc.gen_bits(g["ns"], orc.seed(2, 11), 0, 16))
    rk = dev(orc.gen_range(g["nr"], orc.seed(2, 12), 0, 80))
    rv = dev(orc.gen_bits(g["nr"], orc.seed(2, 13), 0, 16))
    for lf in (0.25, 0.75):
        ht, _ = build(gpu, sk, sw, g["ns"], lf)
        r = probe(gpu, ht, rk, rv, g["nr"])
        assert (r.count, r.sum) == (g["count"], g["sum"])
        ht2, _ = build(gpu, sk, None, g["ns"], lf, mgv.MGV_PAYLOAD_ROWID)
        d = probe(gpu, ht2, rk, None, g["nr"], flags=mgv.MGV_PROBE_DIGEST)
        assert (d.digest_sum, d.digest_xor) == (g["digest_sum"], g["digest_xor"])


def test_join_match_multiset_materialized(gpu, orc):
    import npref
    rng = np.random.default_rng(11)
    ns, nr = 5000, 30011
    sk = rng.integers(-200, 3000, ns).astype(np.int32)
    rk = rng.integers(-300, 3300, nr).astype(np.int32)
    ht, _ = build(gpu, dev(sk), None, ns, 0.5, mgv.MGV_PAYLOAD_ROWID)
    r_idx, s_idx = npref.join_pairs(sk, rk)
    cap = len(r_idx) + 10
    out_r, out_p = mgv.DeviceBuffer(4 * cap), mgv.DeviceBuffer(4 * cap)
    res = mgv.ResultBuffer()
    drk = dev(rk)
    gpu.call("mgv_hashjoin_probe_materialize", ht.buf.vp, ht.nbuckets, drk.vp, nr, 0, out_r.vp, out_p.vp, cap,
             res.vp, mgv.launch(0, 256), None)
    gpu_sync()
    r = res.read()
    assert r.count == len(r_idx) and r.overflow == 0
    got = np.stack([out_r.download(np.int32, r.count), out_p.download(np.int32, r.count)], 1).astype(np.int64)
    want = np.stack([r_idx, s_idx], 1)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(got, want)


def test_join_edge_cases(gpu, orc):
    # negative / extreme keys, empty probe, empty build, reserved slot pattern
    sk = np.array([INT_MIN, -1, 0, 1, INT_MAX, -1, 7], dtype=np.int32)
    sw = np.array([1, 2, 3, 4, 5, -1, 6], dtype=np.int32)   # (-1,-1) is the reserved empty pattern
    rk = np.array([INT_MIN, -1, -1, 0, 2, INT_MAX, 7, 7], dtype=np.int32)
    rv = np.arange(8, dtype=np.int32)
    ht, br = build(gpu, dev(sk), dev(sw), len(sk))
    assert br.errors == 1
    r = probe(gpu, ht, dev(rk), dev(rv), len(rk))
    keep = ~((sk == -1) & (sw == -1))
    t = orc.table(sk[keep].copy(), sw[keep].copy())
    want = orc.probe(t, rk, rv)
    orc.free(t)
    assert (r.count, r.sum) == (want["count"], want["sum"])
    assert probe(gpu, ht, dev(rk), dev(rv), 0).count == 0
    empty, _ = build(gpu, dev(sk), dev(sw), 0)
    assert probe(gpu, empty, dev(rk), dev(rv), len(rk)).count == 0


def test_select_probe_sum_matches_oracle(gpu, orc):
    rng = np.random.default_rng(5)
    ns, nr = 20000, 200003
    sk = orc.gen_perm(ns, ns, orc.seed(4, 3), 9)
    sw = orc.gen_bits(ns, orc.seed(4, 4), 9, 16)
    rk = orc.gen_range(nr, orc.seed(4, 1), 9, 2 * ns)
    rv = orc.gen_bits(nr, orc.seed(4, 2), 9, 16)
    c = orc.gen_bits(nr, orc.seed(4, 0), 9, 31)
    ht, _ = build(gpu, dev(sk), dev(sw), ns)
    t = orc.table(sk, sw)
    dk, dv, dc = dev(rk), dev(rv), dev(c)
    for bound in (INT_MIN, 2**31 // 20, 2**30, INT_MAX):
        want = orc.probe(t, rk, rv, pred=c, bound=bound)
        for compact in (0, 1):
            for block in (64, 1024):
                r = probe(gpu, ht, dk, dv, nr, pred=dc, bound=bound, launch=mgv.launch(0, block, compact))
                assert (r.count, r.sum) == (want["count"], want["sum"]), (bound, compact, block)
    orc.free(t)


# ----------------------------------------------------------------- group-by

def test_select_join_groupby_c3_small(gpu, orc, golden):
    g = golden["c3_small"]
    T = orc.zipf_cdf(g["nranks"], 1.0)
    c = dev(orc.gen_bits(g["nr"], orc.seed(3, 0), 0, 31))
    rk = dev(orc.gen_zipf(g["nr"], orc.seed(3, 1), 0, T, orc.seed(3, 5), 0))
    rv = dev(orc.gen_bits(g["nr"], orc.seed(3, 2), 0, 16))
    sk = dev(orc.gen_perm(g["nranks"], g["nranks"], orc.seed(3, 3), 0))
    sg = dev(orc.gen_bits(g["nranks"], orc.seed(3, 4), 0, 6))
    ht, _ = build(gpu, sk, sg, g["nranks"])
    G = g["ngroups"]
    for case in g["cases"]:
        for compact in (0, 1):
            gc, gs = mgv.DeviceBuffer(8 * G), mgv.DeviceBuffer(8 * G)
            gpu.call("mgv_fill_u64", gc.vp, G, 0, None)
            gpu.call("mgv_fill_u64", gs.vp, G, 0, None)
            gpu.call("mgv_hashjoin_probe_groupby", ht.buf.vp, ht.nbuckets, c.vp, case["bound"], rk.vp, rv.vp,
                     g["nr"], gc.vp, gs.vp, G, mgv.launch(0, 256, compact), None)
            gpu_sync()
            assert gc.download(np.uint64, G).tolist() == case["gcnt"]
            assert gs.download(np.int64, G).tolist() == case["gsum"]


@pytest.mark.parametrize("ngroups", [64, 4096, 1 << 16])
def test_groupby_matches_oracle(gpu, orc, ngroups):
    n = 300007
    T = orc.zipf_cdf(ngroups, 0.8)
    k = orc.gen_zipf(n, orc.seed(5, 0), 0, T)
    k[::1001] = ngroups + 3                     # out-of-range keys are counted as errors
    v = orc.gen_bits(n, orc.seed(5, 1), 0, 16)
    gc, gs = mgv.DeviceBuffer(8 * ngroups), mgv.DeviceBuffer(8 * ngroups)
    gpu.call("mgv_fill_u64", gc.vp, ngroups, 0, None)
    gpu.call("mgv_fill_u64", gs.vp, ngroups, 0, None)
    res = mgv.ResultBuffer()
    dk, dv = dev(k), dev(v)
    gpu.call("mgv_groupby", dk.vp, dv.vp, n, gc.vp, gs.vp, ngroups, res.vp, mgv.launch(0, 256), None)
    gpu_sync()
    wc, ws = orc.groupby(k, v, ngroups)
    assert np.array_equal(gc.download(np.uint64, ngroups), wc)
    assert np.array_equal(gs.download(np.int64, ngroups), ws)
    assert res.read().errors == int((k >= ngroups).sum())


def test_groupby_golden(gpu, orc, golden):
    g = golden["groupby_small"]
    T = orc.zipf_cdf(g["ngroups"], 0.8)
    kz = mgv.DeviceBuffer(4 * g["n"])
    Tp = C.c_void_p()
    gpu.call("mgv_zipf_table_create", C.byref(Tp), g["ngroups"], 0.8)
    gpu.call("mgv_gen_zipf", kz.vp, g["n"], orc.seed(5, 0), 0, Tp, g["ngroups"], 0, 0, None)
    v = gen(gpu, "bits", g["n"], orc.seed(5, 1), 0, 16)
    G = g["ngroups"]
    gc, gs = mgv.DeviceBuffer(8 * G), mgv.DeviceBuffer(8 * G)
    gpu.call("mgv_fill_u64", gc.vp, G, 0, None)
    gpu.call("mgv_fill_u64", gs.vp, G, 0, None)
    gpu.call("mgv_groupby", kz.vp, v.vp, g["n"], gc.vp, gs.vp, G, None, mgv.launch(0, 256), None)
    gpu_sync()
    assert col_check(gc.download(np.uint64, G).astype(np.int64)) == g["gcnt_hash"]
    assert col_check(gs.download(np.int64, G)) == g["gsum_hash"]
    gpu.call("mgv_free", Tp)


# ----------------------------------------------------------------- mid-size exact join

def test_join_mid_size_exact(gpu, orc):
    """C2 distribution at 2^20 build / 2^23 probe, against the chained-table
    oracle (multi-threaded)."""
    ns, nr = 1 << 20, 1 << 23
    sk = gen(gpu, "perm", ns, ns, orc.seed(2, 0), 0)
    sw = gen(gpu, "bits", ns, orc.seed(2, 1), 0, 16)
    rk = gen(gpu, "bits", nr, orc.seed(2, 2), 0, 21)
    rv = gen(gpu, "bits", nr, orc.seed(2, 3), 0, 16)
    ht, br = build(gpu, sk, sw, ns)
    r = probe(gpu, ht, rk, rv, nr, launch=mgv.launch(0, 256))
    hk, hw, hrk, hrv = down(sk, ns), down(sw, ns), down(rk, nr), down(rv, nr)
    assert r.count == int((hrk < ns).sum())
    t = orc.table(hk, hw, log_buckets=21)
    want = orc.probe(t, hrk, hrv, threads=8)
    orc.free(t)
    assert (r.count, r.sum) == (want["count"], want["sum"])


# ----------------------------------------------------------------- partitioned probe

def probe_part(gpu, ht, keys, vals, n, fanout, pred=None, bound=0, flags=0, block=256):
    ncols = 3 if flags & mgv.MGV_PROBE_DIGEST else 2
    ws = mgv.partition_workspace(n, fanout, ncols)
    res = mgv.ResultBuffer()
    gpu.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, pred.vp if pred is not None else None, bound,
             keys.vp, vals.vp if vals is not None else None, n, 0, flags, fanout, ws.vp, ws.nbytes, res.vp,
             mgv.launch(0, block), None)
    gpu_sync()
    return res.read()


@pytest.mark.parametrize("fanout", [2, 8, 64, 256, 512])
def test_partitioned_probe_c2_small(gpu, orc, golden, fanout):
    g = golden["c2_small"]
    sk, sw, rk, rv = c2_small_cols(gpu, orc, g)
    ht, _ = build(gpu, sk, sw, g["ns"])
    r = probe_part(gpu, ht, rk, rv, g["nr"], fanout)
    assert (r.count, r.sum) == (g["count"], g["sum"])
    ht2, _ = build(gpu, sk, None, g["ns"], 0.5, mgv.MGV_PAYLOAD_ROWID)
    d = probe_part(gpu, ht2, rk, None, g["nr"], fanout, flags=mgv.MGV_PROBE_DIGEST)
    assert (d.count, d.digest_sum, d.digest_xor) == (g["count"], g["digest_sum"], g["digest_xor"])


@pytest.mark.parametrize("fanout", [16, 128])
def test_partitioned_select_probe_matches_oracle(gpu, orc, fanout):
    ns, nr = 30000, 500009
    sk = orc.gen_perm(ns, ns, orc.seed(4, 3), 1)
    sw = orc.gen_bits(ns, orc.seed(4, 4), 1, 16)
    rk = orc.gen_range(nr, orc.seed(4, 1), 1, 2 * ns)
    rv = orc.gen_bits(nr, orc.seed(4, 2), 1, 16)
    c = orc.gen_bits(nr, orc.seed(4, 0), 1, 31)
    ht, _ = build(gpu, dev(sk), dev(sw), ns)
    t = orc.table(sk, sw)
    dk, dv, dc = dev(rk), dev(rv), dev(c)
    for bound in (INT_MIN, 2**31 // 100, 2**30, INT_MAX):
        want = orc.probe(t, rk, rv, pred=c, bound=bound)
        r = probe_part(gpu, ht, dk, dv, nr, fanout, pred=dc, bound=bound)
        assert (r.count, r.sum) == (want["count"], want["sum"]), bound
    orc.free(t)


def test_partitioned_probe_duplicates_and_skew(gpu, orc, golden):
    g = golden["dup_join"]
    sk = dev(orc.gen_range(g["ns"], orc.seed(2, 10), 0, 64))
    sw = dev(orc.gen_bits(g["ns"], orc.seed(2, 11), 0, 16))
    rk = dev(orc.gen_range(g["nr"], orc.seed(2, 12), 0, 80))
    rv = dev(orc.gen_bits(g["nr"], orc.seed(2, 13), 0, 16))
    ht, _ = build(gpu, sk, sw, g["ns"])
    for fanout in (4, 32):
        r = probe_part(gpu, ht, rk, rv, g["nr"], fanout)
        assert (r.count, r.sum) == (g["count"], g["sum"])


def test_partitioned_groupby_c3_small(gpu, orc, golden):
    g = golden["c3_small"]
    T = orc.zipf_cdf(g["nranks"], 1.0)
    c = dev(orc.gen_bits(g["nr"], orc.seed(3, 0), 0, 31))
    rk = dev(orc.gen_zipf(g["nr"], orc.seed(3, 1), 0, T, orc.seed(3, 5), 0))
    rv = dev(orc.gen_bits(g["nr"], orc.seed(3, 2), 0, 16))
    sk = dev(orc.gen_perm(g["nranks"], g["nranks"], orc.seed(3, 3), 0))
    sg = dev(orc.gen_bits(g["nranks"], orc.seed(3, 4), 0, 6))
    ht, _ = build(gpu, sk, sg, g["nranks"])
    G = g["ngroups"]
    ws = mgv.partition_workspace(g["nr"], 16, 2)
    for case in g["cases"]:
        gc, gs = mgv.DeviceBuffer(8 * G), mgv.DeviceBuffer(8 * G)
        gpu.call("mgv_fill_u64", gc.vp, G, 0, None)
        gpu.call("mgv_fill_u64", gs.vp, G, 0, None)
        gpu.call("mgv_hashjoin_probe_groupby_partitioned", ht.buf.vp, ht.nbuckets, c.vp, case["bound"], rk.vp, rv.vp,
                 g["nr"], gc.vp, gs.vp, G, 16, ws.vp, ws.nbytes, mgv.launch(0, 256), None)
        gpu_sync()
        assert gc.download(np.uint64, G).tolist() == case["gcnt"]
        assert gs.download(np.int64, G).tolist() == case["gsum"]


def test_partitioned_workspace_too_small_is_an_error(gpu):
    ht = mgv.HashTable(100)
    k = dev(np.arange(1000, dtype=np.int32))
    res = mgv.ResultBuffer()
    small = mgv.DeviceBuffer(1024)
    with pytest.raises(mgv.MgvError) as e:
        gpu.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, None, 0, k.vp, None, 1000, 0, 0, 16,
                 small.vp, small.nbytes, res.vp, mgv.launch(0, 256), None)
    assert e.value.code == mgv.E_CAPACITY
    with pytest.raises(mgv.MgvError) as e:
        gpu.call("mgv_hashjoin_probe_partitioned", ht.buf.vp, ht.nbuckets, None, 0, k.vp, None, 1000, 0, 0, 12,
                 small.vp, small.nbytes, res.vp, mgv.launch(0, 256), None)
    assert e.value.code == mgv.E_INVALID


def test_xcc_ids_cover_eight_xcds(gpu):
    """The XCC id the partitioned probe reads (HW_REG_XCC_ID) names all
    eight XCDs of an MI355X, each of which runs a share of a large grid:
    the per-XCD drain order of k_part_probe relies on both."""
    out = mgv.DeviceBuffer(32)
    gpu.call("mgv_fill_u32", out.vp, 8, 0, None)
    gpu.call("mgv_debug_xcc_histogram", out.vp, 4096, None)
    gpu_sync()
    h = out.download(np.uint32, 8)
    assert int(h.sum()) == 4096
    assert int(h.min()) >= 4096 // 32, h.tolist()


def test_join_mid_size_partitioned_exact(gpu, orc):
    ns, nr = 1 << 20, 1 << 23
    sk = gen(gpu, "perm", ns, ns, orc.seed(2, 0), 0)
    sw = gen(gpu, "bits", ns, orc.seed(2, 1), 0, 16)
    rk = gen(gpu, "bits", nr, orc.seed(2, 2), 0, 21)
    rv = gen(gpu, "bits", nr, orc.seed(2, 3), 0, 16)
    ht, _ = build(gpu, sk, sw, ns)
    direct = probe(gpu, ht, rk, rv, nr)
    for fanout in (64, 256):
        r = probe_part(gpu, ht, rk, rv, nr, fanout)
        assert (r.count, r.sum) == (direct.count, direct.sum)
    assert direct.count == int((down(rk, nr) < ns).sum())
