// mgv_groupby.hip -- GROUP BY over many groups: key-range partitioning, then
// aggregation of each key range in an LDS table.
//
// With more groups than fit an LDS table, a direct group-by pays two global
// 64-bit atomics per row, and Zipf-skewed keys serialise on the hot groups.
// Here the rows are first scattered into F = ngroups / G key ranges
// (partition = key >> log2(G), the partitioning kernels of the probe), then
// persistent blocks claim chunks of a range and accumulate them into a
// G-entry LDS table {u32 count, u64 sum} with LDS atomics, flushing it to the
// global arrays (one pair of atomics per touched group) when they move to
// another range and at the end.  Rows whose key is outside [0, ngroups) are
// dropped and counted in res->errors, as in the direct group-by.  There is no
// reference counterpart (aggregation.cu is scalar only).
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_ops.hpp"
#include "mgv_part.hpp"

namespace mgv {

constexpr uint32_t kGroupLog = 13;                 // G = 8192 groups per key range
constexpr uint32_t kGroupsPerRange = 1u << kGroupLog;
constexpr int kAggBlock = 1024;
constexpr uint32_t kAggChunk = 1u << 16;           // rows per claimed chunk

// Fewer groups than this go to the single-pass LDS group-by.
constexpr uint32_t kDirectMaxGroups = 4096;

struct AggArgs {
    uint64_t *gcnt;
    int64_t *gsum;
    uint32_t ngroups;
    uint32_t *next;      // chunk counter (zeroed before the launch)
};

__device__ __forceinline__ void agg_flush(uint32_t range, const uint32_t *scnt, const unsigned long long *ssum,
                                          const AggArgs &a) {
    const uint32_t base = range << kGroupLog;
    for (uint32_t g = threadIdx.x; g < kGroupsPerRange; g += kAggBlock) {
        const uint32_t c = scnt[g];
        if (c) {
            atomicAdd((unsigned long long *)&a.gcnt[base + g], (unsigned long long)c);
            atomicAdd((unsigned long long *)&a.gsum[base + g], ssum[g]);
        }
    }
}

__global__ __launch_bounds__(kAggBlock) void k_range_groupby(PartLayout L, AggArgs a) {
    __shared__ uint32_t scnt[kGroupsPerRange];
    __shared__ unsigned long long ssum[kGroupsPerRange];
    __shared__ uint32_t s_cum[kMaxFanout];    // cumulative chunk counts per range
    __shared__ uint32_t s_item;
    const uint32_t F = L.F;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        uint32_t acc = 0;
        for (uint32_t p = 0; p < F; ++p) {
            acc += (uint32_t)((L.ptotal[p] + kAggChunk - 1) / kAggChunk);
            s_cum[p] = acc;
        }
    }
    for (uint32_t g = t; g < kGroupsPerRange; g += kAggBlock) {
        scnt[g] = 0;
        ssum[g] = 0;
    }
    __syncthreads();
    const uint32_t total = s_cum[F - 1];
    uint32_t cur = 0xffffffffu;
    for (;;) {
        if (t == 0) s_item = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();                      // s_item is rewritten next round
        if (item >= total) break;
        uint32_t lo = 0, hi = F - 1;          // first range whose cumulative count exceeds item
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_cum[mid] > item) hi = mid; else lo = mid + 1;
        }
        const uint32_t p = lo;
        if (p != cur) {
            if (cur != 0xffffffffu) {
                agg_flush(cur, scnt, ssum, a);
                __syncthreads();
                for (uint32_t g = t; g < kGroupsPerRange; g += kAggBlock) {
                    scnt[g] = 0;
                    ssum[g] = 0;
                }
                __syncthreads();
            }
            cur = p;
        }
        const uint32_t first = p ? s_cum[p - 1] : 0;
        const uint64_t rlo = L.pstart[p] + (uint64_t)(item - first) * kAggChunk;
        const uint64_t pend = L.pstart[p] + L.ptotal[p];
        const uint64_t rhi = (rlo + kAggChunk < pend) ? rlo + kAggChunk : pend;
        const uint32_t base = p << kGroupLog;
        const int32_t *K = L.col[0];
        const int32_t *V = L.col[1];
        // rlo is a multiple of 64 rows (partition starts are aligned), so
        // quads never straddle a partition start
        for (uint64_t i0 = rlo + 4ull * t; i0 < rhi; i0 += 4ull * kAggBlock) {
            if (i0 + 3 < rhi) {
                const v4i k = load_stream4(K + i0);
                const v4i v = load_stream4(V + i0);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t g = (uint32_t)q_get(k, j) - base;
                    atomicAdd(&scnt[g], 1u);
                    atomicAdd(&ssum[g], (unsigned long long)(int64_t)q_get(v, j));
                }
            } else {
                for (uint64_t i = i0; i < rhi; ++i) {
                    const uint32_t g = (uint32_t)K[i] - base;
                    atomicAdd(&scnt[g], 1u);
                    atomicAdd(&ssum[g], (unsigned long long)(int64_t)V[i]);
                }
            }
        }
        __syncthreads();
    }
    if (cur != 0xffffffffu) agg_flush(cur, scnt, ssum, a);
}

static uint32_t range_fanout(uint32_t ngroups) {
    uint32_t F = 2;
    while ((uint64_t)F * kGroupsPerRange < ngroups) F <<= 1;
    return F;
}

}  // namespace mgv

using namespace mgv;

extern "C" {

int mgv_groupby_workspace_size(uint64_t n, uint32_t ngroups, uint64_t *bytes) {
    if (!bytes || ngroups == 0) return fail(MGV_E_INVALID, "mgv_groupby_workspace_size: bad argument");
    if (ngroups <= kDirectMaxGroups || n == 0) {
        *bytes = 0;
        return MGV_OK;
    }
    if (ngroups > kMaxFanout * kGroupsPerRange)
        return fail(MGV_E_INVALID, "mgv_groupby_partitioned: more than %u groups", kMaxFanout * kGroupsPerRange);
    uint64_t need = 0;
    int rc = plan_layout(n, range_fanout(ngroups), 2, nullptr, 0, nullptr, &need);
    if (rc) return rc;
    *bytes = need + 256;   // + the chunk counter
    return MGV_OK;
}

int mgv_groupby_partitioned(const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                            uint32_t ngroups, void *ws, uint64_t ws_bytes, mgv_result *res, mgv_launch LC,
                            mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!keys || !vals || !gcnt || !gsum || ngroups == 0)
        return fail(MGV_E_INVALID, "mgv_groupby_partitioned: bad argument");
    if (n == 0) return MGV_OK;
    if (ngroups <= kDirectMaxGroups) return mgv_groupby(keys, vals, n, gcnt, gsum, ngroups, res, LC, st);
    if (ngroups > kMaxFanout * kGroupsPerRange)
        return fail(MGV_E_INVALID, "mgv_groupby_partitioned: more than %u groups", kMaxFanout * kGroupsPerRange);
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "mgv_groupby_partitioned: n must be < 2^31");
    if (!aligned16(keys) || !aligned16(vals)) return fail(MGV_E_ALIGN, "mgv_groupby_partitioned: unaligned column");
    const uint32_t F = range_fanout(ngroups);
    if (!ws || ws_bytes < 256) return fail(MGV_E_CAPACITY, "mgv_groupby_partitioned: workspace too small");
    PartLayout L;
    int rc = plan_layout(n, F, 2, (char *)ws + 256, ws_bytes - 256, &L, nullptr);
    if (rc) return rc;
    hipStream_t s = stream_of(st);
    PartIn in = {nullptr, 0, keys, vals, 0, 0u, kGroupLog, ngroups, res, nullptr};
    if ((rc = partition_rows(L, in, s, scatter_shape(LC)))) return rc;
    uint32_t *next = (uint32_t *)ws;
    if ((rc = mgv_fill_u32(next, 1, 0, (mgv_stream_t)s))) return rc;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(k_range_groupby, kAggBlock);
    const uint32_t grid = LC.grid ? LC.grid : (uint32_t)cached;
    AggArgs a = {gcnt, gsum, ngroups, next};
    hipLaunchKernelGGL(k_range_groupby, dim3(grid), dim3(kAggBlock), 0, s, L, a);
    return check_launch("k_range_groupby");
}

}  // extern "C"
