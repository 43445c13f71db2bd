// mgv_part.hpp -- layout and entry points of the row partitioning shared by
// the partitioned probe (mgv_partition.hip) and the partitioned group-by
// (mgv_groupby.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mgv.h"
#include "mgv_wave.hpp"

namespace mgv {

constexpr int kPartBlock = 256;            // count kernel
constexpr int kTileMax = 8192;             // largest scatter tile (rows per block are a multiple)
constexpr uint32_t kMaxFanout = 512;        // partitioned probe / group-by
constexpr uint32_t kMaxFanoutWide = 2048;   // second level of the two-level partitioning
constexpr uint32_t kPartAlign = 64;         // partition starts aligned to 64 rows (default)
constexpr uint32_t kItemRows = 4096;       // probe work item (one wave)
constexpr uint32_t kMaxPartBlocks = 4096;

struct PartLayout {
    uint32_t F, nblk, ncols;
    uint32_t align;                  // partition starts are multiples of `align` rows
    uint64_t rows_per_blk, n, cap;   // cap = rows per partitioned column
    uint32_t *hist;                  // [nblk][F]
    unsigned long long *ptotal;      // [F]
    unsigned long long *pstart;      // [F + 1]
    unsigned long long *pcursor;     // [F]
    uint32_t *xhead;                 // [8]
    int32_t *col[3];                 // partitioned k, v, rowid
};

// Input of the partitioning kernels.
struct PartIn {
    const int32_t *pred;
    int32_t bound;
    const int32_t *key;
    const int32_t *val;   // may be null (partitioned as zeros)
    int64_t row_base;
    uint32_t want_rowid;  // third partitioned column = row ids
    uint32_t rshift;      // key-range partitioning: partition = key >> rshift ...
    uint32_t rlimit;      // ... for keys in [0, rlimit); 0 = hash partitioning
    mgv_result *res;      // rows dropped for an out-of-range key -> res->errors
    const unsigned long long *nlimit;   // optional device row count: rows >= *nlimit are ignored
    // Hash partitioning only: 0 = keys in, keys out; 1 = keys in, fmix32(key)
    // out (partition by the hash's top bits as with 0); 2 = hashes in and out
    // (partition by the top bits of the value itself).  fmix32 is a
    // bijection, so hashes compare equal exactly when their keys do.
    uint32_t key_hash;
};

// Plans the workspace of a partitioning of n rows into F partitions with
// ncols partitioned columns; with L == nullptr only *need is set.
// F is a power of two in [2, 2048]; align (rows) a power of two <= 64.
int plan_layout(uint64_t n, uint32_t F, uint32_t ncols, void *ws, uint64_t ws_bytes, PartLayout *L, uint64_t *need,
                uint32_t align = kPartAlign);

// Launches prep, count, scan and scatter on stream s; shape is one of the
// MGV_SCATTER_* workgroup shapes of the scatter (0 = 256 threads x 4096 rows).
int partition_rows(const PartLayout &L, const PartIn &in, hipStream_t s, uint32_t shape = 0);

// The scatter shape requested by a launch descriptor (bits 8-11 of its flags).
inline uint32_t scatter_shape(const mgv_launch &L) { return (L.reserved >> 8) & 0xfu; }

// Partition of an already hashed key: the top bits of the hash.
__device__ __forceinline__ uint32_t part_of_hash(int32_t h, uint32_t F) {
    return (uint32_t)(((uint64_t)(uint32_t)h * F) >> 32);
}

// Loads the 4 rows [i0, i0+4) of a column (vector load when all are in range).
__device__ __forceinline__ v4i load_quad(const int32_t *p, uint64_t i0, uint64_t hi) {
    if (i0 + 3 < hi) return load_stream4(p + i0);
    v4i q = {0, 0, 0, 0};
    if (i0 < hi) q.x = load_stream1(p + i0);
    if (i0 + 1 < hi) q.y = load_stream1(p + i0 + 1);
    if (i0 + 2 < hi) q.z = load_stream1(p + i0 + 2);
    return q;
}

// Exclusive scan of one value per thread over a BLOCK-thread block:
// wave-level inclusive scan with ds_bpermute shuffles, then the wave totals.
template <int BLOCK>
__device__ __forceinline__ uint32_t scan_block_excl(uint32_t v, uint32_t *wtot) {
    const uint32_t l = lane_id(), w = wave_id();
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (l >= (uint32_t)d) x += y;
    }
    if (l == kWave - 1) wtot[w] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < w; ++i) before += wtot[i];
    __syncthreads();
    return before + x - v;
}


}  // namespace mgv
