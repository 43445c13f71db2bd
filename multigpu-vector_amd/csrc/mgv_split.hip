// mgv_split.hip -- the split operator: the rows of a vector routed to the
// parents of a split by the reference's partition function.
//
// The reference's split (src/operators/split.cu:54-118) computes, per row,
// dest = hash3(x) = x & 1 of the row's value (49-51), compacts each
// destination's rows per warp with a ballot into a 5*32-row staging area in
// global memory, and hands every full 4*32-row chunk to that parent's
// consume_warp; at_close (131-150) flushes what is left.  Rows whose
// destination has no parent (an odd row with one parent) are dropped, and
// parents beyond the second receive nothing.
//
// Here a split appends each destination's rows to that parent's output
// columns: a block takes a tile of rows (512 threads x 32 rows with one
// column, 1024 x 8 with two to four), every wave ballots its rows per
// destination, the block adds up its waves' counts in LDS and reserves one
// contiguous run per destination with a single global atomic, and each wave
// writes its rows into its part of the run (each ballot's rows to
// consecutive positions).  A vector is read once and each kept row written
// once: 4 * ncols bytes read per row (plus 4 for a filter column) and
// 4 * ncols written per kept row, HBM-bound.  The order of the rows inside a
// destination is unspecified, as it is in the reference.
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_wave.hpp"

namespace mgv {

constexpr int kSplitBlock = 1024;
constexpr int kSplitBlock1 = 512;                              // one column
constexpr int kSplitDests = 2;                                 // x & 1 has two values

// Rows per lane per tile; each tile costs one cursor atomic per destination.
// One column: 32 rows per lane in 512-thread blocks, two per CU (16384-row
// tiles, 128 VGPRs); two to four columns: 16 or 32 rows per lane spill, so 8
// rows per lane in 1024-thread blocks (8192-row tiles).
template <int NC>
constexpr int split_rows() { return NC == 1 ? 32 : 8; }

struct SplitArgs {
    const int32_t *pred;
    int32_t bound;
    const int32_t *col[MGV_SPLIT_MAX_COLS];
    uint64_t n;
    uint32_t ndest;                                   // min(nparents, 2)
    int32_t *out[kSplitDests][MGV_SPLIT_MAX_COLS];
    uint64_t cap;
    unsigned long long *counts;
};

// The second launch bound is HIP's minimum waves per SIMD: 4 for 512-thread
// blocks keeps two blocks (16 waves) per CU within 128 VGPRs.
template <int NC, int BLOCK>
__global__ __launch_bounds__(BLOCK, 2048 / BLOCK) void k_split(SplitArgs a) {
    constexpr int R = split_rows<NC>();
    constexpr uint32_t kWaves = BLOCK / kWave;
    constexpr uint64_t kTile = (uint64_t)BLOCK * R;
    __shared__ uint32_t wcnt[kSplitDests][kWaves];
    __shared__ unsigned long long sbase[kSplitDests];
    const uint32_t lane = lane_id();
    const uint32_t w = wave_id();
    const uint64_t ntiles = (a.n + kTile - 1) / kTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // Tile-uniform base pointers and 32-bit row offsets inside the tile
        // (one scalar base and one offset register per access, not a 64-bit
        // address per row); row (lane, k) of the wave is w*64*R + k*64 + lane.
        const uint64_t first = t * kTile;
        const uint32_t rows = (uint32_t)(a.n - first < kTile ? a.n - first : kTile);
        const uint32_t off0 = w * (kWave * R) + lane;
        int32_t v[NC][R];
        int32_t pv[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t off = off0 + (uint32_t)k * kWave;
            const bool ok = off < rows;
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][k] = ok ? a.col[c][first + off] : 0;
        }
        if (a.pred) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t off = off0 + (uint32_t)k * kWave;
                pv[k] = off < rows ? a.pred[first + off] : 0;
            }
        }
        // Destinations as per-lane bit masks (bit k = row k): hash3
        // (split.cu:49-51) is the low bit of the row's first column; rows
        // outside the vector, filtered out or without a parent are in
        // neither mask.  Plain bit operations here, so no lane mask is shared
        // between the count and the write phase (64 of them would spill).
        uint32_t keep = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const bool ok = off0 + (uint32_t)k * kWave < rows && (!a.pred || pv[k] <= a.bound);
            keep |= (uint32_t)ok << k;
        }
        uint32_t odd = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) odd |= ((uint32_t)v[0][k] & 1u) << k;
        const uint32_t dmask[kSplitDests] = {keep & ~odd, a.ndest > 1 ? keep & odd : 0u};
        const uint64_t tot = wave_sum_u64((uint64_t)__popc(dmask[0]) | ((uint64_t)__popc(dmask[1]) << 32));
        if (lane == 0) {
            wcnt[0][w] = (uint32_t)tot;
            wcnt[1][w] = (uint32_t)(tot >> 32);
        }
        __syncthreads();
        if (threadIdx.x < kSplitDests) {
            const int d = threadIdx.x;
            uint64_t sum = 0;
            for (uint32_t i = 0; i < kWaves; ++i) sum += wcnt[d][i];
            sbase[d] = sum ? atomicAdd(a.counts + d, (unsigned long long)sum) : 0ull;
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < kSplitDests; ++d) {
            uint64_t base = sbase[d];
            for (uint32_t i = 0; i < w; ++i) base += wcnt[d][i];
            // this wave's run starts at `base`; positions >= cap are not written
            const uint64_t room64 = base < a.cap ? a.cap - base : 0;
            const uint32_t room = room64 < 0xffffffffull ? (uint32_t)room64 : 0xffffffffu;
            uint32_t at = 0;                           // rows of this wave's run so far
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const bool mine = (dmask[d] >> k) & 1u;
                const uint64_t b = ballot(mine);
                const uint32_t pos = at + mbcnt(b);
                if (mine && pos < room) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) a.out[d][c][base + pos] = v[c][k];
                }
                at += popc64(b);
            }
        }
        __syncthreads();   // wcnt and sbase are rewritten by the next tile
    }
}

}  // namespace mgv

using namespace mgv;

extern "C" {

int mgv_split(const int32_t *pred, int32_t bound, const int32_t *const *cols, uint32_t ncols, uint64_t n,
              uint32_t nparents, int32_t *const *out, uint64_t cap, uint64_t *counts, mgv_launch L, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!cols || !out || !counts) return fail(MGV_E_INVALID, "mgv_split: null argument");
    if ((uintptr_t)counts & 7u) return fail(MGV_E_ALIGN, "mgv_split: counts must be 8-byte aligned");
    if (ncols == 0 || ncols > MGV_SPLIT_MAX_COLS)
        return fail(MGV_E_INVALID, "mgv_split: ncols must be in [1, %d]", MGV_SPLIT_MAX_COLS);
    if (nparents == 0 || nparents > MGV_SPLIT_MAX_PARENTS)
        return fail(MGV_E_INVALID, "mgv_split: nparents must be in [1, %d]", MGV_SPLIT_MAX_PARENTS);
    SplitArgs a = {};
    a.pred = pred;
    a.bound = bound;
    a.n = n;
    a.ndest = nparents < (uint32_t)kSplitDests ? nparents : (uint32_t)kSplitDests;
    a.cap = cap;
    a.counts = (unsigned long long *)counts;
    for (uint32_t c = 0; c < ncols; ++c) {
        if (!cols[c]) return fail(MGV_E_INVALID, "mgv_split: column %u is null", c);
        a.col[c] = cols[c];
    }
    for (uint32_t d = 0; d < a.ndest; ++d)
        for (uint32_t c = 0; c < ncols; ++c) {
            if (!out[d * ncols + c]) return fail(MGV_E_INVALID, "mgv_split: output column %u of parent %u is null", c, d);
            a.out[d][c] = out[d * ncols + c];
        }
    if (n == 0) return MGV_OK;
    const int block = ncols == 1 ? kSplitBlock1 : kSplitBlock;
    const uint64_t tile = (uint64_t)block * (ncols == 1 ? split_rows<1>() : split_rows<4>());
    const uint64_t ntiles = (n + tile - 1) / tile;
    static thread_local int cached1 = 0, cached = 0;
    if (!cached1) cached1 = persistent_blocks(k_split<1, kSplitBlock1>, kSplitBlock1);
    if (!cached) cached = persistent_blocks(k_split<4, kSplitBlock>, kSplitBlock);
    uint64_t grid = L.grid ? L.grid : (uint64_t)(ncols == 1 ? cached1 : cached);
    if (grid > ntiles) grid = ntiles;
    hipStream_t s = stream_of(st);
    switch (ncols) {
        case 1: hipLaunchKernelGGL((k_split<1, kSplitBlock1>), dim3((uint32_t)grid), dim3(block), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_split<2, kSplitBlock>), dim3((uint32_t)grid), dim3(block), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_split<3, kSplitBlock>), dim3((uint32_t)grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((k_split<4, kSplitBlock>), dim3((uint32_t)grid), dim3(block), 0, s, a); break;
    }
    return check_launch("k_split");
}

}  // extern "C"
