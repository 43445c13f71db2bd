"""Writes tests/golden/golden.json from the independent numpy implementation
(npref.py).  Run:  python tests/golden/make_golden.py

Each case records the generator parameters (seeds/streams/sizes), a checksum
of every generated column, and the expected query results.  The tests
regenerate the columns with the C oracle and the HIP library, check the
column checksums, and compare results with these values.
"""
from __future__ import annotations

import json
import os

import numpy as np

import npref as R

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")


def col_sum(x: np.ndarray) -> dict:
    x = np.asarray(x)
    idx = np.arange(len(x), dtype=np.uint64)
    h = R.mix64((idx << np.uint64(32)) | x.astype(np.uint32).astype(np.uint64))
    with np.errstate(over="ignore"):
        hs = int(np.sum(h, dtype=np.uint64))
    return {"n": int(len(x)), "sum": int(x.astype(np.int64).sum()), "hash": hs}


def main() -> None:
    g: dict = {"about": "written by tests/golden/make_golden.py from tests/golden/npref.py"}

    # Philox words at assorted counters.
    g["words"] = []
    for (cfg, col, stream, idx) in [(1, 0, 0, 0), (1, 1, 0, 1), (2, 3, 7, 12345), (4, 2, 255, 999999), (0, 0, 0, 3)]:
        s = R.seed(cfg, col)
        w = R.words(idx + 1, s, stream)[idx]
        g["words"].append({"seed": s, "stream": stream, "i": idx, "word": int(w)})

    # fmix32 values.
    xs = [0, 1, 2, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 123456789]
    g["fmix32"] = [{"x": x, "h": int(R.fmix32(x))} for x in xs]

    # Feistel permutations on small ranges.
    g["perm"] = []
    for rng, cfg, col, stream in [(16, 2, 0, 0), (37, 4, 3, 5), (1000, 3, 3, 0), (4096, 2, 0, 0)]:
        s = R.seed(cfg, col)
        p = R.perm(np.arange(rng), rng, s, stream)
        g["perm"].append({"range": rng, "seed": s, "stream": stream, "head": [int(v) for v in p[:16]],
                          "check": col_sum(p)})

    # Zipf thresholds.
    g["zipf"] = []
    for n, s in [(16, 1.0), (16, 0.8), (1024, 1.0), (65536, 0.8)]:
        T = R.zipf_cdf(n, s)
        g["zipf"].append({"n": n, "s": s, "head": [int(t) for t in T[:8]], "sum": int(T.astype(np.uint64).sum()),
                          "mid": int(T[n // 2])})

    # C1-shaped filtered sum (small): a in [0,2^20), b in [0,2^31), b <= 2^30-1.
    n = 1 << 14 | 77
    a = R.gen_bits(n, R.seed(1, 0), 0, 20)
    b = R.gen_bits(n, R.seed(1, 1), 0, 31)
    bound = (1 << 30) - 1
    cnt, s64, s32 = R.select_sum(b, a, bound)
    g["c1_small"] = {"n": n, "a": col_sum(a), "b": col_sum(b), "bound": bound, "count": cnt, "sum64": s64,
                     "sum32": s32}

    # The reference's own single-column form (main9.cu:56-61): values in [1, 10000], <= 5000.
    x = R.gen_range(n, R.seed(0, 0), 0, 10000) + 1
    cnt, s64, s32 = R.select_sum(x, x, 5000)
    g["ref_sum_select"] = {"n": n, "x": col_sum(x), "bound": 5000, "count": cnt, "sum64": s64, "sum32": s32}
    # Self-join (main9.cu:63-69): count = sum_v hist[v]^2 over v <= 5000.
    hist = np.bincount(x[x <= 5000], minlength=10001)
    r_idx, s_idx = R.join_pairs(x[x <= 5000], x[x <= 5000])
    assert len(r_idx) == int((hist.astype(np.int64) ** 2).sum())
    g["ref_selfjoin"] = {"n": n, "bound": 5000, "count": int((hist.astype(np.int64) ** 2).sum())}

    # C2-shaped join (small): S.k = perm of [0, 2^12), R.k in [0, 2^13).
    ns, nr = 1 << 12, (1 << 14) + 5
    sk = R.perm(np.arange(ns), ns, R.seed(2, 0), 0)
    sw = R.gen_bits(ns, R.seed(2, 1), 0, 16)
    rk = R.gen_bits(nr, R.seed(2, 2), 0, 13)
    rv = R.gen_bits(nr, R.seed(2, 3), 0, 16)
    res = R.join_agg(sk, sw, rk, rv)
    assert res["count"] == int((rk < ns).sum())
    dig = R.join_agg(sk, np.arange(ns, dtype=np.int32), rk, None)
    g["c2_small"] = {"ns": ns, "nr": nr, "sk": col_sum(sk), "sw": col_sum(sw), "rk": col_sum(rk),
                     "rv": col_sum(rv), "count": res["count"], "sum": res["sum"],
                     "digest_sum": dig["digest_sum"], "digest_xor": dig["digest_xor"]}

    # Duplicate build keys: S.k in [0, 64) (many duplicates), R.k in [0, 80).
    ns2, nr2 = 3000, 5000
    sk2 = R.gen_range(ns2, R.seed(2, 10), 0, 64)
    sw2 = R.gen_bits(ns2, R.seed(2, 11), 0, 16)
    rk2 = R.gen_range(nr2, R.seed(2, 12), 0, 80)
    rv2 = R.gen_bits(nr2, R.seed(2, 13), 0, 16)
    res2 = R.join_agg(sk2, sw2, rk2, rv2)
    dig2 = R.join_agg(sk2, np.arange(ns2, dtype=np.int32), rk2, None)
    g["dup_join"] = {"ns": ns2, "nr": nr2, "sk": col_sum(sk2), "rk": col_sum(rk2), "count": res2["count"],
                     "sum": res2["sum"], "digest_sum": dig2["digest_sum"], "digest_xor": dig2["digest_xor"]}

    # C3-shaped select -> join -> group-by (small), Zipf(1.0) probe keys.
    nranks, nr3, ngroups = 1 << 10, (1 << 15) + 3, 64
    T = R.zipf_cdf(nranks, 1.0)
    c = R.gen_bits(nr3, R.seed(3, 0), 0, 31)
    rk3 = R.gen_zipf(nr3, R.seed(3, 1), 0, T, R.seed(3, 5), 0)
    rv3 = R.gen_bits(nr3, R.seed(3, 2), 0, 16)
    sk3 = R.perm(np.arange(nranks), nranks, R.seed(3, 3), 0)
    sg3 = R.gen_bits(nranks, R.seed(3, 4), 0, 6)
    cases = []
    for t in [21474836, 214748365, 1073741824, 1932735283]:
        gc, gs = R.join_groupby(sk3, sg3, rk3, rv3, c, t - 1, ngroups)
        agg = R.join_agg(sk3, None, rk3, rv3, c, t - 1)
        cases.append({"t": t, "bound": t - 1, "gcnt": [int(v) for v in gc], "gsum": [int(v) for v in gs],
                      "count": agg["count"], "sum": agg["sum"]})
    g["c3_small"] = {"nranks": nranks, "nr": nr3, "ngroups": ngroups, "c": col_sum(c), "rk": col_sum(rk3),
                     "rv": col_sum(rv3), "sk": col_sum(sk3), "sg": col_sum(sg3), "cases": cases}

    # Plain group-by with Zipf(0.8) keys over 2^12 groups.
    ng, n5 = 1 << 12, 20000
    T5 = R.zipf_cdf(ng, 0.8)
    k5 = R.gen_zipf(n5, R.seed(5, 0), 0, T5)
    v5 = R.gen_bits(n5, R.seed(5, 1), 0, 16)
    gc5, gs5 = R.groupby(k5, v5, ng)
    g["groupby_small"] = {"ngroups": ng, "n": n5, "k": col_sum(k5), "v": col_sum(v5),
                          "gcnt_hash": col_sum(gc5.astype(np.int64)), "gsum_hash": col_sum(gs5),
                          "total_sum": int(gs5.sum())}

    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
