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

#include <vector>

#include "mgv_internal.hpp"
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

}  // extern "C"
