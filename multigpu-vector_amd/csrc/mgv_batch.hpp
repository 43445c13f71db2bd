// mgv_batch.hpp -- the partition stage shared by the batch probes that
// scatter their rows first (the XCD-sliced probe in mgv_fused.hip, the LDS
// batch probe in mgv_probe_lds.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mgv.h"
#include "mgv_ops.hpp"

namespace mgv {

constexpr int kRegionWords = 4;        // per query: first region row, rows per region, partitions F, slice buckets
constexpr int kBatchMaxParts = 256;    // partitions (regions) per query at most
constexpr int kBatchTile = 4096;       // rows per scatter item (up to 32 partitions a query)
constexpr int kBatchTileWide = 16384;  // rows per scatter item (up to 256 partitions a query)

// Rows of the batch queries partitioned into regions: query q's rows with
// partition p (= top bits of fmix32(key): (h * F) >> 32) go to rows
// [regions[4q] + p * cap, ... + cap) of the region columns, cap =
// regions[4q + 1], F = regions[4q + 2] (0: the query is not partitioned).
struct BatchScatter {
    const uint64_t *regions;
    unsigned long long *fill;   // per query kBatchMaxParts words: rows reserved in each region (may pass cap)
    int32_t *rkey, *rval;       // region columns
    uint32_t *ctr;              // claim counter of the scatter items (zeroed before the launch)
    uint64_t nscatter;          // scatter items: up to `tile` rows of one partitioned query each
    uint32_t hashed;            // 1: the key column receives fmix32(key)
    uint32_t tile;              // kBatchTile or kBatchTileWide
    // rows of BUILD queries past a full region: appended here and inserted
    // into their tables after the slices are written back (null: no builds)
    unsigned long long *ocount;
    uint32_t *oquery;
    int32_t *okey, *opay;
};

// Rows per region of a query of n rows in F regions: n / F plus eight
// standard deviations of a uniform hash partition's size, plus 256, in
// multiples of 64 (rows past a full region are probed on the spot).
inline uint64_t batch_region_cap(uint64_t n, uint32_t F) {
    const double lam = (double)n / F;
    const uint64_t c = (uint64_t)(lam + 8.0 * __builtin_sqrt(lam)) + 256;
    return (c + 63) / 64 * 64;
}

// Workspace of a partitioned batch: 256 bytes of claim counters, the fill
// counts (kBatchMaxParts per query), the two region columns.
inline uint64_t batch_fill_bytes(uint32_t nq) { return ((uint64_t)nq * kBatchMaxParts * 8 + 255) / 256 * 256; }
inline uint64_t batch_ws_bytes(uint32_t nq, uint64_t region_rows) {
    const uint64_t col = (region_rows * 4 + 255) / 256 * 256;
    return 256 + batch_fill_bytes(nq) + 2 * col;
}

// The LDS batch build's overflow list: query index, key and payload per row.
inline uint64_t lds_batch_ovf_bytes(uint64_t build_rows) { return 3 * ((build_rows * 4 + 255) / 256 * 256); }

// Launches the scatter (k_batch_scatter) of every scatter item.
int launch_batch_scatter(const mgv_query *queries, const mgv_work_item *items, const BatchScatter &S, hipStream_t s);

// Probes one row against a global table, following its chain: count and
// SUM(val + payload) of every hitting slot (the empty pattern never hits).
__device__ __forceinline__ void probe_row(const uint64_t *slots, uint32_t nb, uint32_t key, int32_t val,
                                          uint64_t &cnt, uint64_t &sum) {
    uint32_t b = bucket_of((int32_t)key, nb);
    for (uint32_t pass = 0; pass < nb; ++pass) {
        const v4u *g = reinterpret_cast<const v4u *>(slots + (uint64_t)b * 8);
        v4u q[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) q[w] = load_global(g + w);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const uint32_t k = (t & 1) ? q[t >> 1].z : q[t >> 1].x;
            const uint32_t p = (t & 1) ? q[t >> 1].w : q[t >> 1].y;
            if (k == key && !(k == ~0u && p == ~0u)) {
                cnt += 1;
                sum += (uint64_t)((int64_t)(int32_t)p + (int64_t)val);
            }
        }
        if (q[3].z == ~0u && q[3].w == ~0u) break;   // a bucket with an empty slot ends the chain
        b = (b + 1 == nb) ? 0 : b + 1;
    }
}

}  // namespace mgv
