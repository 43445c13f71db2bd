// mgv_fused.hip -- fused multi-query persistent kernel.
//
// The reference pushes every vector of every query through its own kernel
// launch (launch_consume_pipeline, src/operators/exchange.cu:124-194) between
// per-operator open/close launches with device synchronisations
// (src/operators/operator.cu:126-159).  Here a batch of queries is described
// by a device table of mgv_query records and a flat list of work items (one
// vector = a row range of one query).  One persistent launch drains the list:
// each wave claims an item with one atomic, runs that query's operator chain
// over the item's rows, and flushes its partial result with wave-level
// atomics, so waves of different queries run side by side with no block
// barrier in the loop.  Builds and probes of the same batch are two launches
// on one stream (normally captured together in one hipGraph).
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "mgv_batch.hpp"
#include "mgv_internal.hpp"
#include "mgv_part.hpp"
#include "mgv_pipes.hpp"

namespace mgv {

// NC: the chain reads columns [0, NC) (see load_full in mgv_pipes.hpp).
template <int NC, class Op>
__device__ __forceinline__ void run_rows(Op &op, const Cols &cols, uint64_t lo, uint64_t hi, int64_t row_base) {
    op.consume_open();
    drive_chunks<NC>(op, cols, lo, kChunk, hi, row_base);
    op.consume_close();
}

template <int W, bool COMPACT, bool DIGEST>
__device__ __forceinline__ void run_probe_sum(const mgv_query &q, uint64_t lo, uint64_t hi, void *lds) {
    typedef OpAggWave<2, DIGEST> Agg;
    if (q.pred) {
        OpSelect<OpProbe<Agg>, COMPACT, DIGEST ? 3 : 2, W> op;
        op.sm = reinterpret_cast<typename decltype(op)::Smem *>(lds);
        op.bound = q.bound;
        op.parent.slots = q.slots;
        op.parent.nbuckets = q.nbuckets;
        op.parent.parent.res = q.res;
        Cols cols = {{q.pred, q.key, q.val, nullptr}, DIGEST ? 3 : -1};
        run_rows<3>(op, cols, lo, hi, q.row_base);
    } else {
        OpProbe<Agg> op;
        op.slots = q.slots;
        op.nbuckets = q.nbuckets;
        op.parent.res = q.res;
        Cols cols = {{q.key, q.val, nullptr, nullptr}, DIGEST ? 2 : -1};
        run_rows<2>(op, cols, lo, hi, q.row_base);
    }
}

template <int W, bool COMPACT>
__device__ __forceinline__ void run_probe_group(const mgv_query &q, uint64_t lo, uint64_t hi, void *lds) {
    if (q.pred) {
        OpSelect<OpProbe<OpGroupWave>, COMPACT, 2, W> op;
        op.sm = reinterpret_cast<typename decltype(op)::Smem *>(lds);
        op.bound = q.bound;
        op.parent.slots = q.slots;
        op.parent.nbuckets = q.nbuckets;
        auto &g = op.parent.parent;
        g.gcnt = q.gcnt;
        g.gsum = q.gsum;
        g.ngroups = q.ngroups;
        g.res = q.res;
        Cols cols = {{q.pred, q.key, q.val, nullptr}, -1};
        run_rows<3>(op, cols, lo, hi, q.row_base);
    } else {
        OpProbe<OpGroupWave> op;
        op.slots = q.slots;
        op.nbuckets = q.nbuckets;
        op.parent.gcnt = q.gcnt;
        op.parent.gsum = q.gsum;
        op.parent.ngroups = q.ngroups;
        op.parent.res = q.res;
        Cols cols = {{q.key, q.val, nullptr, nullptr}, -1};
        run_rows<2>(op, cols, lo, hi, q.row_base);
    }
}

template <int W>
__device__ __forceinline__ void run_build(const mgv_query &q, uint64_t lo, uint64_t hi, void *lds) {
    const bool rowid = q.payload_mode == MGV_PAYLOAD_ROWID;
    const int32_t *pay = q.payload_mode == MGV_PAYLOAD_COLUMN ? q.val : nullptr;
    if (q.pred) {
        OpSelect<OpBuild<W>, true, 2, W> op;
        op.sm = reinterpret_cast<typename decltype(op)::Smem *>(lds);
        op.bound = q.bound;
        op.parent.slots = q.slots;
        op.parent.nbuckets = q.nbuckets;
        op.parent.res = q.res;
        Cols cols = {{q.pred, q.key, pay, nullptr}, rowid ? 2 : -1};
        run_rows<3>(op, cols, lo, hi, q.row_base);
    } else {
        OpBuild<W> op;
        op.slots = q.slots;
        op.nbuckets = q.nbuckets;
        op.res = q.res;
        Cols cols = {{q.key, pay, nullptr, nullptr}, rowid ? 1 : -1};
        run_rows<2>(op, cols, lo, hi, q.row_base);
    }
}

template <int W>
__device__ __forceinline__ void run_select_sum(const mgv_query &q, uint64_t lo, uint64_t hi) {
    OpSelect<OpAggWave<1, false>, false, 1, W> op;
    op.sm = nullptr;   // predication: no stage
    op.bound = q.pred ? q.bound : 0x7fffffff;
    op.parent.res = q.res;
    Cols cols = {{q.pred ? q.pred : q.val, q.val, nullptr, nullptr}, -1};
    run_rows<2>(op, cols, lo, hi, q.row_base);
}

__device__ __forceinline__ void run_groupby(const mgv_query &q, uint64_t lo, uint64_t hi) {
    OpGroupWave op;
    op.gcnt = q.gcnt;
    op.gsum = q.gsum;
    op.ngroups = q.ngroups;
    op.res = q.res;
    Cols cols = {{q.key, q.val, nullptr, nullptr}, -1};
    run_rows<2>(op, cols, lo, hi, q.row_base);
}

template <int BLOCK, bool COMPACT>
__global__ __launch_bounds__(BLOCK) void k_fused(const mgv_query *__restrict__ Q, const mgv_work_item *__restrict__ items,
                                                 uint64_t nitems, uint32_t *counter) {
    constexpr int W = BLOCK / kWave;
    // compaction stages of every wave: 3 columns x 320 rows (the largest chain)
    __shared__ __align__(16) int32_t lds[W * 3 * 320 + 16];
    static_assert(sizeof(typename OpSelect<OpProbe<OpAggWave<2, true>>, true, 3, W>::Smem) <= sizeof(lds), "lds");
    for (;;) {
        uint32_t it = 0;
        if (lane_id() == 0) it = atomicAdd(counter, 1u);
        it = __builtin_amdgcn_readlane(it, 0);
        if (it >= nitems) break;
        const mgv_work_item wi = items[it];
        const mgv_query q = Q[wi.query];
        const uint64_t lo = wi.lo;
        const uint64_t hi = (wi.lo + wi.rows < q.n) ? wi.lo + wi.rows : q.n;
        switch (q.kind) {
            case MGV_Q_SELECT_SUM: run_select_sum<W>(q, lo, hi); break;
            case MGV_Q_PROBE_SUM:
                if (q.flags & MGV_PROBE_DIGEST) run_probe_sum<W, COMPACT, true>(q, lo, hi, lds);
                else run_probe_sum<W, COMPACT, false>(q, lo, hi, lds);
                break;
            case MGV_Q_PROBE_GROUPBY: run_probe_group<W, COMPACT>(q, lo, hi, lds); break;
            case MGV_Q_GROUPBY: run_groupby(q, lo, hi); break;
            case MGV_Q_BUILD: run_build<W>(q, lo, hi, lds); break;
            default: break;
        }
    }
}

template <int B, bool C>
static int fused_b(const mgv_query *Q, const mgv_work_item *items, uint64_t nitems, uint32_t *counter,
                   const mgv_launch &L, hipStream_t s) {
    auto kern = k_fused<B, C>;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(kern, B);
    uint64_t grid = L.grid ? L.grid : (uint64_t)cached;
    const uint64_t need = (nitems + (B / kWave) - 1) / (B / kWave);
    if (need < grid) grid = need ? need : 1;
    hipLaunchKernelGGL(kern, dim3((uint32_t)grid), dim3(B), 0, s, Q, items, nitems, counter);
    return check_launch("k_fused");
}

template <bool C>
static int fused_c(const mgv_query *Q, const mgv_work_item *items, uint64_t nitems, uint32_t *counter,
                   const mgv_launch &L, hipStream_t s) {
    switch (L.block) {
        case 64: return fused_b<64, C>(Q, items, nitems, counter, L, s);
        case 128: return fused_b<128, C>(Q, items, nitems, counter, L, s);
        case 0:
        case 256: return fused_b<256, C>(Q, items, nitems, counter, L, s);
        case 512: return fused_b<512, C>(Q, items, nitems, counter, L, s);
        case 1024: return fused_b<1024, C>(Q, items, nitems, counter, L, s);
        default: return fail(MGV_E_INVALID, "block %u not in {64,128,256,512,1024}", L.block);
    }
}


// ------------------------------------------------------------------ XCD-sliced probe
//
// A probe's random bucket reads that miss the L2 cost one 128-byte line each:
// on MI355X about 55 G lines/s whatever the table size beyond the L2 (a
// 64-byte read over 32 MiB to 1 GiB: 55-67 G/s), while an XCD's 4 MB L2
// serves about 210 G reads/s.  When every XCD probes every table, each L2
// must hold the whole table being probed; when the waves of XCD x probe only
// the rows whose home bucket lies in the x-th eighth of the table, the eight
// L2s hold one eighth each, and random reads over a 32 MiB table run at
// 208 G/s instead of 67 (measured).  The sliced batch probe therefore
// scatters the selected rows of each large PROBE_SUM query by the top bits
// of fmix32(key) -- the bits that pick the bucket's 8th, 16th or 32nd of
// the table -- into 8, 16 or 32 regions (k_batch_scatter); then one
// persistent launch (k_slice_probe) lets the waves of XCD x take the regions
// r = x mod 8 of every query first, region after region, then whatever is
// left of the other XCDs' regions, and the other queries' vectors follow in
// a k_fused launch.  The slice an XCD probes at a time stays in its L2
// beside the streamed rows, and no table line is cached by two XCDs.

constexpr int kSliceParts = 8;                  // XCDs: sections of the probe launch
constexpr uint32_t kItemQueryBits = 24;         // work item .query: query | region << 24

struct SlicedRun {
    const uint64_t *regions;     // per query kRegionWords words (mgv_batch.hpp); F = 0: not sliced
    unsigned long long *fill;    // per query kBatchMaxParts words, one per region: rows reserved (may pass the capacity)
    int32_t *rkey, *rval;        // region columns
    uint32_t *ctr;               // claim counters: [0] scatter items, [1 + s] probe section s
    uint64_t sec[kSliceParts + 2];   // probe items of section s: [sec[s], sec[s + 1])
};

// Scatter items: up to kBatchTile rows of one partitioned query.  The block
// selects the rows (pred <= bound), sorts them in LDS by partition (top bits
// of fmix32(key)), reserves each partition's run with one atomic, and
// writes {key or fmix32(key), val} runs; rows past a full region's capacity
// (skewed keys) are probed on the spot and added to the query's result.
template <int BLOCK, int T>
__global__ __launch_bounds__(BLOCK) void k_batch_scatter(const mgv_query *__restrict__ Q,
                                                         const mgv_work_item *__restrict__ items, BatchScatter R) {
    constexpr int NQ = T / (4 * BLOCK);
    constexpr int PER = kBatchMaxParts / BLOCK > 0 ? kBatchMaxParts / BLOCK : 1;
    static_assert(NQ >= 1 && T % (4 * BLOCK) == 0 && (BLOCK >= kBatchMaxParts || kBatchMaxParts % BLOCK == 0),
                  "batch scatter shape");
    __shared__ int32_t sk[T];
    __shared__ int32_t sv[T];
    __shared__ uint32_t cnt[kBatchMaxParts], start[kBatchMaxParts + 1], lim[kBatchMaxParts];
    __shared__ unsigned long long dst[kBatchMaxParts];
    __shared__ uint32_t wtot[BLOCK / kWave];
    __shared__ uint32_t s_item;
    const uint32_t t = threadIdx.x;
    // a tile's rows: key and value quads, and the mask of the rows in range
    // that pass the filter
    v4i k[NQ], v[NQ];
    uint32_t m[NQ];
    auto load_tile = [&](const mgv_work_item &wi, const mgv_query &q) {
        const uint64_t hi = (wi.lo + wi.rows < q.n) ? wi.lo + wi.rows : q.n;
        // a BUILD's payload column only with MGV_PAYLOAD_COLUMN (NONE: zeros)
        const bool payload = q.val && (q.kind != MGV_Q_BUILD || q.payload_mode == MGV_PAYLOAD_COLUMN);
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const uint64_t i0 = wi.lo + 4ull * t + (uint64_t)j * 4 * BLOCK;
            uint32_t mm = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) mm |= (uint32_t)(i0 + e < hi) << e;
            k[j] = load_quad(q.key, i0, hi);
            v[j] = payload ? load_quad(q.val, i0, hi) : v4i{0, 0, 0, 0};
            if (q.pred) {
                const v4i c = load_quad(q.pred, i0, hi);
#pragma unroll
                for (int e = 0; e < 4; ++e) mm &= ~((uint32_t)(q_get(c, e) > q.bound) << e);
            }
            m[j] = mm;
        }
    };
    if (t == 0) s_item = atomicAdd(&R.ctr[0], 1u);
    for (uint32_t p = t; p < (uint32_t)kBatchMaxParts; p += BLOCK) cnt[p] = 0;
    __syncthreads();
    uint32_t it = s_item;
    mgv_work_item wi = {};
    mgv_query q = {};
    if (it < R.nscatter) {
        wi = items[it];
        q = Q[wi.query];
        load_tile(wi, q);
    }
    // the next tile's rows load while this tile's runs are written
    while (it < R.nscatter) {
        const uint64_t base = R.regions[kRegionWords * wi.query], cap = R.regions[kRegionWords * wi.query + 1];
        const uint32_t F = (uint32_t)R.regions[kRegionWords * wi.query + 2];
        const uint32_t qi = wi.query;
        __syncthreads();   // every thread has read s_item
        if (t == 0) s_item = atomicAdd(&R.ctr[0], 1u);   // the next tile
        uint32_t rank[NQ][4];
#pragma unroll
        for (int j = 0; j < NQ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                rank[j][e] = ((m[j] >> e) & 1u)
                                 ? atomicAdd(&cnt[part_of_hash((int32_t)fmix32((uint32_t)q_get(k[j], e)), F)], 1u)
                                 : 0u;
        __syncthreads();
        // partition starts (block scan over F counts) and run reservations
        uint32_t c[PER], mine = 0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t p = t * PER + e;
            c[e] = p < F ? cnt[p] : 0u;
            mine += c[e];
        }
        uint32_t st = scan_block_excl<BLOCK>(mine, wtot);
        if (t == BLOCK - 1) start[kBatchMaxParts] = st + mine;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t p = t * PER + e;
            if (p < F) {
                start[p] = st;
                uint32_t fit = 0;
                if (c[e]) {
                    const unsigned long long off = atomic_add_global(
                        &R.fill[(uint64_t)qi * kBatchMaxParts + p], (unsigned long long)c[e]);
                    const uint64_t room = off < cap ? cap - off : 0;
                    fit = room < c[e] ? (uint32_t)room : c[e];
                    dst[p] = base + (uint64_t)p * cap + off;
                }
                lim[p] = fit;
                st += c[e];
            }
        }
        __syncthreads();
        const uint32_t nit = s_item;
        for (uint32_t p = t; p < (uint32_t)kBatchMaxParts; p += BLOCK) cnt[p] = 0;   // read before the barrier above
#pragma unroll
        for (int j = 0; j < NQ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if ((m[j] >> e) & 1u) {
                    const uint32_t h = fmix32((uint32_t)q_get(k[j], e));
                    const uint32_t pos = start[part_of_hash((int32_t)h, F)] + rank[j][e];
                    sk[pos] = R.hashed ? (int32_t)h : q_get(k[j], e);
                    sv[pos] = q_get(v[j], e);
                }
        const mgv_query qc = q;
        if (nit < R.nscatter) {
            wi = items[nit];
            q = Q[wi.query];
            load_tile(wi, q);
        }
        __syncthreads();
        const uint32_t total = start[kBatchMaxParts];
        uint64_t ocnt = 0, osum = 0;
        bool over = false;
        for (uint32_t i = t; i < total; i += BLOCK) {
            const int32_t x = sk[i], val = sv[i];
            const uint32_t h = R.hashed ? (uint32_t)x : fmix32((uint32_t)x);
            const uint32_t p = part_of_hash((int32_t)h, F);
            const uint32_t j = i - start[p];
            if (j < lim[p]) {
                R.rkey[dst[p] + j] = x;
                R.rval[dst[p] + j] = val;
            } else if (qc.kind == MGV_Q_BUILD) {
                const unsigned long long o = atomic_add_global(R.ocount, 1ull);
                R.oquery[o] = qi;
                R.okey[o] = (int32_t)(R.hashed ? fmix32_inv(h) : (uint32_t)x);
                R.opay[o] = val;
            } else {
                probe_row(qc.slots, qc.nbuckets, R.hashed ? fmix32_inv(h) : (uint32_t)x, val, ocnt, osum);
                over = true;
            }
        }
        if (ballot(over)) {
            const uint64_t a = wave_sum_u64(ocnt), b2 = wave_sum_u64(osum);
            if (lane_id() == 0 && a) {
                atomic_add_global((unsigned long long *)&qc.res->count, (unsigned long long)a);
                atomic_add_global((unsigned long long *)&qc.res->sum, (unsigned long long)b2);
            }
        }
        it = nit;
    }
}

template <int BLOCK, int T>
static int batch_scatter_k(const mgv_query *queries, const mgv_work_item *items, const BatchScatter &S,
                           hipStream_t s) {
    auto kern = k_batch_scatter<BLOCK, T>;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(kern, BLOCK);
    const uint64_t grid = S.nscatter < (uint64_t)cached ? S.nscatter : (uint64_t)cached;
    hipLaunchKernelGGL(kern, dim3((uint32_t)grid), dim3(BLOCK), 0, s, queries, items, S);
    return check_launch("k_batch_scatter");
}

// Items of kBatchTile rows: 256-thread blocks; of kBatchTileWide rows (up
// to 256 partitions a tile: runs four times longer): 1024-thread blocks.
int launch_batch_scatter(const mgv_query *queries, const mgv_work_item *items, const BatchScatter &S,
                         hipStream_t s) {
    if (!S.nscatter) return MGV_OK;
    if (S.tile == (uint32_t)kBatchTileWide) return batch_scatter_k<1024, kBatchTileWide>(queries, items, S, s);
    if (S.tile == (uint32_t)kBatchTile) return batch_scatter_k<256, kBatchTile>(queries, items, S, s);
    return fail(MGV_E_INVALID, "batch scatter: tile of %u rows", S.tile);
}

// The region probe of the sliced batch: a wave on XCD x claims items of
// section x (the regions r = x mod 8 of every sliced query, query by query),
// then of the other sections.  An item's .query holds the query index and,
// above bit 24, its region.  (Section 8, the other queries' vectors, runs as
// a k_fused launch of its own: its select-and-compact chain needs far more
// registers, which would halve this kernel's waves.)
template <int BLOCK, bool COOP>
__global__ __launch_bounds__(BLOCK) void k_slice_probe(const mgv_query *__restrict__ Q,
                                                       const mgv_work_item *__restrict__ items, SlicedRun R) {
    const uint32_t x = xcc_id();
    uint32_t step = 0;   // position in the section order x, x + 1, ..., x + 7 (wave-uniform)
    for (;;) {
        uint32_t s = 0, it = 0;
        bool got = false;
        while (step < kSliceParts) {
            s = (x + step) & (kSliceParts - 1);
            const uint64_t len = R.sec[s + 1] - R.sec[s];
            uint32_t c = 0;
            if (lane_id() == 0) c = atomicAdd(&R.ctr[1 + s], 1u);
            c = __builtin_amdgcn_readlane(c, 0);
            if (c < len) {
                it = c;
                got = true;
                break;
            }
            ++step;
        }
        if (!got) break;
        const mgv_work_item wi = items[R.sec[s] + it];
        const uint32_t qi = wi.query & ((1u << kItemQueryBits) - 1u), region = wi.query >> kItemQueryBits;
        const mgv_query q = Q[qi];
        {
            const uint64_t base = R.regions[kRegionWords * qi], cap = R.regions[kRegionWords * qi + 1];
            const unsigned long long f = R.fill[(uint64_t)qi * kBatchMaxParts + region];
            const uint64_t n = f < cap ? f : cap;
            const uint64_t hi = (wi.lo + wi.rows < n) ? wi.lo + wi.rows : n;
            if (wi.lo < hi) {
                typename std::conditional<COOP, OpProbeCoop<OpAggWave<2, false>>, OpProbe<OpAggWave<2, false>>>::type
                    op;
                op.slots = q.slots;
                op.nbuckets = q.nbuckets;
                op.parent.res = q.res;
                const uint64_t r0 = base + (uint64_t)region * cap;
                Cols cols = {{R.rkey + r0, R.rval + r0, nullptr, nullptr}, -1};
                run_rows<2>(op, cols, wi.lo, hi, 0);
            }
        }
    }
}

template <int B, bool COOP>
static int slice_probe_k(const mgv_query *Q, const mgv_work_item *items, const SlicedRun &R, const mgv_launch &L,
                         hipStream_t s) {
    auto kern = k_slice_probe<B, COOP>;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(kern, B);
    const uint64_t nitems = R.sec[kSliceParts] - R.sec[0];
    uint64_t grid = L.grid ? L.grid : (uint64_t)cached;
    const uint64_t need = (nitems + (B / kWave) - 1) / (B / kWave);
    if (need < grid) grid = need ? need : 1;
    hipLaunchKernelGGL(kern, dim3((uint32_t)grid), dim3(B), 0, s, Q, items, R);
    return check_launch("k_slice_probe");
}

// MGV_LAUNCH_COOP_PROBE: the regions' rows through the 4-lanes-per-row
// bucket scan (one 16-byte load per lane and row: a quarter of the
// texture-addresser work of one row per lane, which bounds the probe once
// the buckets come from the L2)
template <int B>
static int slice_probe_b(const mgv_query *Q, const mgv_work_item *items, const SlicedRun &R, const mgv_launch &L,
                         hipStream_t s) {
    if (L.reserved & MGV_LAUNCH_COOP_PROBE) return slice_probe_k<B, true>(Q, items, R, L, s);
    return slice_probe_k<B, false>(Q, items, R, L, s);
}

static int slice_probe(const mgv_query *Q, const mgv_work_item *items, const SlicedRun &R, const mgv_launch &L,
                       hipStream_t s) {
    switch (L.block) {
        case 64: return slice_probe_b<64>(Q, items, R, L, s);
        case 128: return slice_probe_b<128>(Q, items, R, L, s);
        case 0:
        case 256: return slice_probe_b<256>(Q, items, R, L, s);
        case 512: return slice_probe_b<512>(Q, items, R, L, s);
        case 1024: return slice_probe_b<1024>(Q, items, R, L, s);
        default: return fail(MGV_E_INVALID, "block %u not in {64,128,256,512,1024}", L.block);
    }
}

}  // namespace mgv

using namespace mgv;

extern "C" {

int mgv_batch_plan(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, mgv_work_item *items,
                   uint64_t cap, uint64_t *nitems) {
    if (!queries || !nitems) return fail(MGV_E_INVALID, "mgv_batch_plan: null argument");
    if (vector_rows == 0 || vector_rows % kChunk)
        return fail(MGV_E_INVALID, "mgv_batch_plan: vector_rows %u must be a positive multiple of %d", vector_rows,
                    kChunk);
    uint64_t k = 0;
    for (uint32_t qi = 0; qi < nqueries; ++qi) {
        const mgv_query &q = queries[qi];
        if (q.kind < MGV_Q_SELECT_SUM || q.kind > MGV_Q_BUILD)
            return fail(MGV_E_INVALID, "mgv_batch_plan: query %u has bad kind %u", qi, q.kind);
        for (const int32_t *p : {q.pred, q.key, q.val})
            if (p && !aligned16(p)) return fail(MGV_E_ALIGN, "mgv_batch_plan: query %u column not 16-byte aligned", qi);
        if ((q.kind == MGV_Q_PROBE_SUM || q.kind == MGV_Q_PROBE_GROUPBY || q.kind == MGV_Q_BUILD) &&
            (!q.slots || !q.nbuckets || !q.key))
            return fail(MGV_E_INVALID, "mgv_batch_plan: query %u needs a table and keys", qi);
        if (q.kind == MGV_Q_BUILD && q.n >= (uint64_t)q.nbuckets * MGV_SLOTS_PER_BUCKET)
            return fail(MGV_E_CAPACITY, "mgv_batch_plan: query %u build does not fit its table", qi);
        if ((q.kind == MGV_Q_PROBE_GROUPBY || q.kind == MGV_Q_GROUPBY) && (!q.gcnt || !q.gsum || !q.ngroups))
            return fail(MGV_E_INVALID, "mgv_batch_plan: query %u needs group arrays", qi);
        if ((q.kind == MGV_Q_SELECT_SUM || q.kind == MGV_Q_PROBE_SUM) && !q.res)
            return fail(MGV_E_INVALID, "mgv_batch_plan: query %u needs a result", qi);
        if (q.kind == MGV_Q_SELECT_SUM && !q.val) return fail(MGV_E_INVALID, "mgv_batch_plan: query %u needs val", qi);
        for (uint64_t lo = 0; lo < q.n; lo += vector_rows, ++k) {
            if (items && k < cap) {
                items[k].query = qi;
                items[k].rows = (uint32_t)((q.n - lo < vector_rows) ? q.n - lo : vector_rows);
                items[k].lo = lo;
            }
        }
    }
    *nitems = k;
    if (items && k > cap) return fail(MGV_E_CAPACITY, "mgv_batch_plan: %llu items > capacity %llu",
                                      (unsigned long long)k, (unsigned long long)cap);
    return MGV_OK;
}

int mgv_batch_run(const mgv_query *queries, const mgv_work_item *items, uint64_t nitems, uint32_t *counter,
                  mgv_launch L, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!queries || !counter || (nitems && !items)) return fail(MGV_E_INVALID, "mgv_batch_run: null argument");
    if (nitems == 0) return MGV_OK;
    if (nitems >= 0xffffffffull) return fail(MGV_E_INVALID, "mgv_batch_run: too many items");
    hipStream_t s = stream_of(st);
    int rc = mgv_fill_u32(counter, 1, 0, (mgv_stream_t)s);
    if (rc) return rc;
    if (L.compact == 0) return fused_c<false>(queries, items, nitems, counter, L, s);
    return fused_c<true>(queries, items, nitems, counter, L, s);
}

int mgv_batch_plan_sliced(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, uint64_t min_rows,
                          uint32_t nregions, mgv_work_item *items, uint64_t cap, uint64_t *regions,
                          mgv_sliced_plan *plan) {
    if (!queries || !plan) return fail(MGV_E_INVALID, "mgv_batch_plan_sliced: null argument");
    if (nregions == 0) nregions = 32;
    if (nregions != 8 && nregions != 16 && nregions != 32)
        return fail(MGV_E_INVALID, "mgv_batch_plan_sliced: %u regions per query (8, 16 or 32)", nregions);
    if (nqueries >= (1u << kItemQueryBits)) return fail(MGV_E_INVALID, "mgv_batch_plan_sliced: too many queries");
    uint64_t n = 0;
    int rc = mgv_batch_plan(queries, nqueries, vector_rows, nullptr, 0, &n);   // validates the queries
    if (rc) return rc;
    auto sliced = [&](const mgv_query &q) {
        return q.kind == MGV_Q_PROBE_SUM && !(q.flags & MGV_PROBE_DIGEST) && q.n > 0 && q.n >= min_rows;
    };
    uint64_t k = 0, rows = 0;
    auto put = [&](uint32_t qi, uint64_t lo, uint64_t r) {
        if (items && k < cap) {
            items[k].query = qi;
            items[k].rows = (uint32_t)r;
            items[k].lo = lo;
        }
        ++k;
    };
    // regions
    for (uint32_t qi = 0; qi < nqueries; ++qi) {
        const bool sl = sliced(queries[qi]);
        const uint64_t c = sl ? batch_region_cap(queries[qi].n, nregions) : 0;
        if (regions) {
            regions[kRegionWords * qi] = rows;
            regions[kRegionWords * qi + 1] = c;
            regions[kRegionWords * qi + 2] = sl ? nregions : 0;
            regions[kRegionWords * qi + 3] = 0;
        }
        rows += (uint64_t)nregions * c;
    }
    // scatter tiles of the sliced queries
    for (uint32_t qi = 0; qi < nqueries; ++qi)
        if (sliced(queries[qi]))
            for (uint64_t lo = 0; lo < queries[qi].n; lo += kBatchTile)
                put(qi, lo, queries[qi].n - lo < (uint64_t)kBatchTile ? queries[qi].n - lo : kBatchTile);
    plan->nscatter = k;
    // probe sections x = 0..7: regions x, x + 8, ... of every sliced query,
    // query by query (items over each region's capacity)
    for (int x = 0; x < kSliceParts; ++x) {
        plan->section[x] = k;
        for (uint32_t qi = 0; qi < nqueries; ++qi)
            if (sliced(queries[qi])) {
                const uint64_t c = batch_region_cap(queries[qi].n, nregions);
                for (uint32_t r = (uint32_t)x; r < nregions; r += kSliceParts)
                    for (uint64_t lo = 0; lo < c; lo += vector_rows)
                        put(qi | (r << kItemQueryBits), lo, c - lo < vector_rows ? c - lo : vector_rows);
            }
    }
    // section 8: every other query, in vectors
    plan->section[kSliceParts] = k;
    for (uint32_t qi = 0; qi < nqueries; ++qi)
        if (!sliced(queries[qi]))
            for (uint64_t lo = 0; lo < queries[qi].n; lo += vector_rows)
                put(qi, lo, queries[qi].n - lo < vector_rows ? queries[qi].n - lo : vector_rows);
    plan->section[kSliceParts + 1] = k;
    plan->nitems = k;
    plan->nqueries = nqueries;
    plan->nregions = nregions;
    plan->region_rows = rows;
    plan->workspace_bytes = batch_ws_bytes(nqueries, rows);
    if (items && k > cap) return fail(MGV_E_CAPACITY, "mgv_batch_plan_sliced: %llu items > capacity %llu",
                                      (unsigned long long)k, (unsigned long long)cap);
    return MGV_OK;
}

int mgv_batch_run_sliced(const mgv_query *queries, const mgv_work_item *items, const uint64_t *regions,
                         const mgv_sliced_plan *plan, void *ws, uint64_t ws_bytes, mgv_launch L, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!queries || !plan || !regions || !ws || (plan->nitems && !items))
        return fail(MGV_E_INVALID, "mgv_batch_run_sliced: null argument");
    if (plan->nitems >= 0xffffffffull) return fail(MGV_E_INVALID, "mgv_batch_run_sliced: too many items");
    if (plan->nscatter > plan->section[0] || plan->section[kSliceParts + 1] != plan->nitems ||
        (plan->nregions != 8 && plan->nregions != 16 && plan->nregions != 32))
        return fail(MGV_E_INVALID, "mgv_batch_run_sliced: inconsistent plan");
    for (int p = 0; p <= kSliceParts; ++p)
        if (plan->section[p] > plan->section[p + 1]) return fail(MGV_E_INVALID, "mgv_batch_run_sliced: inconsistent plan");
    const uint64_t need = batch_ws_bytes(plan->nqueries, plan->region_rows);
    if (ws_bytes < need) return fail(MGV_E_CAPACITY, "mgv_batch_run_sliced: workspace %llu < %llu bytes",
                                     (unsigned long long)ws_bytes, (unsigned long long)need);
    if (!aligned16(ws)) return fail(MGV_E_ALIGN, "mgv_batch_run_sliced: workspace not 16-byte aligned");
    if (plan->nitems == 0) return MGV_OK;
    hipStream_t s = stream_of(st);
    char *w = (char *)ws;
    const uint64_t fills = ((uint64_t)plan->nqueries * kBatchMaxParts * 8 + 255) / 256 * 256;
    const uint64_t col = (plan->region_rows * 4 + 255) / 256 * 256;
    SlicedRun R = {};
    R.regions = regions;
    R.ctr = (uint32_t *)w;
    R.fill = (unsigned long long *)(w + 256);
    R.rkey = (int32_t *)(w + 256 + fills);
    R.rval = (int32_t *)(w + 256 + fills + col);
    for (int p = 0; p <= kSliceParts + 1; ++p) R.sec[p] = plan->section[p];
    // counters and fill counts: a fill kernel (the call can be captured into a graph)
    int rc = mgv_fill_u64((uint64_t *)ws, (256 + fills) / 8, 0, (mgv_stream_t)s);
    if (rc) return rc;
    const BatchScatter S = {regions, R.fill, R.rkey, R.rval, R.ctr, plan->nscatter, 0u, (uint32_t)kBatchTile,
                            nullptr, nullptr, nullptr, nullptr};
    if ((rc = launch_batch_scatter(queries, items, S, s))) return rc;
    if (plan->section[kSliceParts] > plan->section[0] && (rc = slice_probe(queries, items, R, L, s))) return rc;
    // section 8: the other queries' vectors, as mgv_batch_run
    const uint64_t nd = plan->section[kSliceParts + 1] - plan->section[kSliceParts];
    if (!nd) return MGV_OK;
    const mgv_work_item *di = items + plan->section[kSliceParts];
    uint32_t *dc = R.ctr + 1 + kSliceParts;
    if (L.compact == 0) return fused_c<false>(queries, di, nd, dc, L, s);
    return fused_c<true>(queries, di, nd, dc, L, s);
}

}  // extern "C"
