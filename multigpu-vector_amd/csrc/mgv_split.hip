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
// columns: one block of 1024 threads takes a tile of 8192 rows, every wave
// ballots its rows per destination, the block adds up its waves' counts in
// LDS and reserves one contiguous run per destination with a single global
// atomic, and each wave writes its rows into its part of the run (each
// ballot's rows to consecutive positions).  A vector is read once and each
// kept row written once: 4 * ncols bytes read per row (plus 4 for a filter
// column) and 4 * ncols written per kept row, HBM-bound.  The order of the
// rows inside a destination is unspecified, as it is in the reference.
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_wave.hpp"

namespace mgv {

constexpr int kSplitBlock = 1024;
constexpr int kSplitRows = 8;                                  // rows per lane per tile
constexpr uint32_t kSplitWaves = kSplitBlock / kWave;
constexpr uint64_t kSplitTile = (uint64_t)kSplitBlock * kSplitRows;   // 8192 rows
constexpr int kSplitDests = 2;                                 // x & 1 has two values

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

template <int NC>
__global__ __launch_bounds__(kSplitBlock) void k_split(SplitArgs a) {
    __shared__ uint32_t wcnt[kSplitDests][kSplitWaves];
    __shared__ unsigned long long sbase[kSplitDests];
    const uint32_t lane = lane_id();
    const uint32_t w = wave_id();
    const uint64_t ntiles = (a.n + kSplitTile - 1) / kSplitTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // row of (lane, k): consecutive lanes read consecutive rows
        const uint64_t r0 = t * kSplitTile + (uint64_t)w * (kWave * kSplitRows) + lane;
        int32_t v[NC][kSplitRows];
        int dest[kSplitRows];
#pragma unroll
        for (int k = 0; k < kSplitRows; ++k) {
            const uint64_t r = r0 + (uint64_t)k * kWave;
            const bool ok = r < a.n;
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][k] = ok ? a.col[c][r] : 0;
            bool keep = ok;
            if (a.pred) keep = keep && (ok ? a.pred[r] : 0) <= a.bound;
            const int d = v[0][k] & 1;            // hash3 (split.cu:49-51)
            dest[k] = (keep && (uint32_t)d < a.ndest) ? d : -1;
        }
        uint64_t b[kSplitDests][kSplitRows];
#pragma unroll
        for (int d = 0; d < kSplitDests; ++d) {
            uint32_t tot = 0;
#pragma unroll
            for (int k = 0; k < kSplitRows; ++k) {
                b[d][k] = ballot(dest[k] == d);
                tot += popc64(b[d][k]);
            }
            if (lane == 0) wcnt[d][w] = tot;
        }
        __syncthreads();
        if (threadIdx.x < kSplitDests) {
            const int d = threadIdx.x;
            uint64_t sum = 0;
            for (uint32_t i = 0; i < kSplitWaves; ++i) sum += wcnt[d][i];
            sbase[d] = sum ? atomicAdd(a.counts + d, (unsigned long long)sum) : 0ull;
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < kSplitDests; ++d) {
            uint64_t base = sbase[d];
            for (uint32_t i = 0; i < w; ++i) base += wcnt[d][i];
#pragma unroll
            for (int k = 0; k < kSplitRows; ++k) {
                const uint64_t pos = base + mbcnt(b[d][k]);
                if (dest[k] == d && pos < a.cap) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) a.out[d][c][pos] = v[c][k];
                }
                base += popc64(b[d][k]);
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
    const uint64_t ntiles = (n + kSplitTile - 1) / kSplitTile;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(k_split<4>, kSplitBlock);
    uint64_t grid = L.grid ? L.grid : (uint64_t)cached;
    if (grid > ntiles) grid = ntiles;
    hipStream_t s = stream_of(st);
    switch (ncols) {
        case 1: hipLaunchKernelGGL(k_split<1>, dim3((uint32_t)grid), dim3(kSplitBlock), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_split<2>, dim3((uint32_t)grid), dim3(kSplitBlock), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_split<3>, dim3((uint32_t)grid), dim3(kSplitBlock), 0, s, a); break;
        default: hipLaunchKernelGGL(k_split<4>, dim3((uint32_t)grid), dim3(kSplitBlock), 0, s, a); break;
    }
    return check_launch("k_split");
}

}  // extern "C"
