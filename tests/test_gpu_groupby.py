"""GPU tests of the key-range partitioned GROUP BY (mgv_groupby_partitioned)
against the oracle's direct-address group-by and against the direct GPU
group-by: per-group COUNT and int64 SUM arrays must be identical, rows with
out-of-range keys are dropped and counted."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mgv

pytestmark = pytest.mark.gpu


def dev(a):
    return mgv.device_array(np.ascontiguousarray(a, dtype=np.int32))


def run_partitioned(keys, vals, ngroups, block=1024, grid=0):
    n = len(keys)
    dk, dv = dev(keys), dev(vals)
    g = mgv.DeviceBuffer(16 * ngroups)
    mgv.call("mgv_fill_u64", g.vp, 2 * ngroups, 0, None)
    need = C.c_uint64()
    mgv.call("mgv_groupby_workspace_size", n, ngroups, C.byref(need))
    ws = mgv.DeviceBuffer(max(256, need.value))
    res = mgv.ResultBuffer()
    mgv.call("mgv_groupby_partitioned", dk.vp, dv.vp, n, g.vp, g.at(8 * ngroups), ngroups, ws.vp, ws.nbytes,
             res.vp, mgv.launch(grid, block), None)
    mgv.sync()
    cnt = g.download(np.uint64, ngroups)
    s = g.download(np.int64, ngroups, 8 * ngroups)
    return cnt, s, res.read().errors


@pytest.mark.parametrize("n,ngroups,zipf", [(1, 1 << 16, 0.8), (63, 5000, 0.0), (1000003, 1 << 16, 0.8),
                                            (300001, 1 << 20, 1.0), (2000000, 8193, 0.0)])
def test_partitioned_groupby_matches_oracle(gpu, orc, n, ngroups, zipf):
    if zipf > 0:
        T = orc.zipf_cdf(ngroups, zipf)
        k = orc.gen_zipf(n, orc.seed(5, 0), 3, T, orc.seed(5, 1), 3)
    else:
        k = orc.gen_range(n, orc.seed(5, 0), 3, ngroups)
    v = orc.gen_bits(n, orc.seed(5, 2), 3, 31) - (1 << 30)     # signed values
    cnt, s, errs = run_partitioned(k, v, ngroups)
    wc, ws = orc.groupby(k, v, ngroups)
    assert errs == 0
    assert np.array_equal(cnt, wc) and np.array_equal(s, ws)


def test_partitioned_groupby_drops_out_of_range_keys(gpu, orc):
    n, ngroups = 500000, 40000
    k = orc.gen_range(n, orc.seed(5, 3), 0, ngroups + 3000) - 1000     # [-1000, ngroups + 2000)
    v = orc.gen_bits(n, orc.seed(5, 4), 0, 16)
    cnt, s, errs = run_partitioned(k, v, ngroups)
    ok = (k >= 0) & (k < ngroups)
    wc, ws = orc.groupby(np.ascontiguousarray(k[ok]), np.ascontiguousarray(v[ok]), ngroups)
    assert errs == int((~ok).sum())
    assert np.array_equal(cnt, wc) and np.array_equal(s, ws)


@pytest.mark.parametrize("block,grid", [(1024, 0), (1024, 7)])
def test_partitioned_equals_direct_groupby(gpu, orc, block, grid):
    n, ngroups = 3 << 20, 1 << 16
    T = orc.zipf_cdf(ngroups, 0.8)
    k = orc.gen_zipf(n, orc.seed(5, 5), 1, T, orc.seed(5, 6), 1)
    v = orc.gen_bits(n, orc.seed(5, 7), 1, 16)
    cnt, s, _ = run_partitioned(k, v, ngroups, block, grid)
    dk, dv = dev(k), dev(v)
    g = mgv.DeviceBuffer(16 * ngroups)
    mgv.call("mgv_fill_u64", g.vp, 2 * ngroups, 0, None)
    res = mgv.ResultBuffer()
    mgv.call("mgv_groupby", dk.vp, dv.vp, n, g.vp, g.at(8 * ngroups), ngroups, res.vp, mgv.launch(0, 256), None)
    mgv.sync()
    assert np.array_equal(cnt, g.download(np.uint64, ngroups))
    assert np.array_equal(s, g.download(np.int64, ngroups, 8 * ngroups))
    assert int(cnt.sum()) == n


def test_partitioned_groupby_small_group_count_uses_direct_path(gpu, orc):
    n, ngroups = 100000, 1000
    k = orc.gen_range(n, orc.seed(5, 8), 0, ngroups)
    v = orc.gen_bits(n, orc.seed(5, 9), 0, 16)
    need = C.c_uint64()
    mgv.call("mgv_groupby_workspace_size", n, ngroups, C.byref(need))
    assert need.value == 0
    cnt, s, errs = run_partitioned(k, v, ngroups)
    wc, ws = orc.groupby(k, v, ngroups)
    assert errs == 0 and np.array_equal(cnt, wc) and np.array_equal(s, ws)
