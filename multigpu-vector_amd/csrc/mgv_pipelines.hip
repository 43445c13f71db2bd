// mgv_pipelines.hip -- the per-vector / resident pipeline driver and the
// select, hash-join and aggregation entry points of the C ABI.
//
// k_pipeline<Pipe, BLOCK> is the MI355X counterpart of launch_consume_pipeline
// (src/operators/exchange.cu:124-194): every block runs consume_open, its
// waves walk the rows in 256-row wave chunks with a grid stride, then
// consume_close.  Launched once per vector it is the reference's plain
// per-vector path; launched once over a resident column with a persistent grid
// it is the single-pass form.
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_pipes.hpp"

namespace mgv {

// LDS-privatised group tables: 1024 groups (12 KiB) or 4096 groups (48 KiB).
constexpr uint32_t kLdsGroupsSmall = 1024;
constexpr uint32_t kLdsGroups = 4096;

}  // namespace mgv

using namespace mgv;

#define MGV_REQUIRE_INIT()                                                            \
    do {                                                                              \
        if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called"); \
    } while (0)

extern "C" {

int mgv_result_clear(mgv_result *r, uint32_t n, mgv_stream_t s) {
    if (!r) return fail(MGV_E_INVALID, "null result");
    return mgv_fill_u64((uint64_t *)r, (uint64_t)n * (sizeof(mgv_result) / 8), 0, s);
}

int mgv_ht_plan(uint64_t nkeys, double load_factor, uint32_t *nbuckets, uint64_t *bytes) {
    if (!(load_factor > 0.0 && load_factor <= 0.9)) return fail(MGV_E_INVALID, "load factor must be in (0, 0.9]");
    double slots = (double)(nkeys ? nkeys : 1) / load_factor;
    uint64_t nb = (uint64_t)((slots + MGV_SLOTS_PER_BUCKET - 1) / MGV_SLOTS_PER_BUCKET);
    if (nb < 1) nb = 1;
    while (nb * MGV_SLOTS_PER_BUCKET <= nkeys) ++nb;   // keep at least one empty slot
    // Large tables get a multiple of 256 buckets so that every partition of a
    // partitioned probe (fan-out <= 256) maps to a whole range of buckets.
    if (nb >= 4096) nb = (nb + 255) / 256 * 256;
    if (nb > 0xffffffffull) return fail(MGV_E_INVALID, "table too large");
    if (nbuckets) *nbuckets = (uint32_t)nb;
    if (bytes) *bytes = nb * MGV_SLOTS_PER_BUCKET * sizeof(uint64_t);
    return MGV_OK;
}

int mgv_ht_clear(uint64_t *slots, uint32_t nbuckets, mgv_stream_t s) {
    return mgv_fill_u64(slots, (uint64_t)nbuckets * MGV_SLOTS_PER_BUCKET, ~0ull, s);
}

int mgv_select_sum(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_result *res,
                   mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!pred || !val || !res) return fail(MGV_E_INVALID, "mgv_select_sum: null argument");
    if (n == 0) return MGV_OK;
    Cols cols = {{pred, val, nullptr, nullptr}, -1};
    if (L.compact == 1) {
        PSelectSum<true>::Params p = {bound, res};
        return launch_pipe<PSelectSum<true>>(p, cols, n, 0, L, s, "select_sum");
    }
    PSelectSum<false>::Params p = {bound, res};
    return launch_pipe<PSelectSum<false>>(p, cols, n, 0, L, s, "select_sum");
}

int mgv_select_materialize(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, int32_t *out,
                           uint64_t cap, mgv_result *res, mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!pred || !val || !res || (!out && cap)) return fail(MGV_E_INVALID, "mgv_select_materialize: null argument");
    if (n == 0) return MGV_OK;
    Cols cols = {{pred, val, nullptr, nullptr}, -1};
    PSelectMat::Params p = {bound, out, cap, res};
    return launch_pipe<PSelectMat>(p, cols, n, 0, L, s, "select_materialize");
}

int mgv_groupby(const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                uint32_t ngroups, mgv_result *res, mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!keys || !vals || !gcnt || !gsum || ngroups == 0) return fail(MGV_E_INVALID, "mgv_groupby: bad argument");
    if (n == 0) return MGV_OK;
    Cols cols = {{keys, vals, nullptr, nullptr}, -1};
    if (ngroups <= kLdsGroupsSmall) {
        PGroupBy<kLdsGroupsSmall>::Params p = {gcnt, gsum, ngroups, res};
        return launch_pipe<PGroupBy<kLdsGroupsSmall>>(p, cols, n, 0, L, s, "groupby");
    }
    if (ngroups <= kLdsGroups) {
        PGroupBy<kLdsGroups>::Params p = {gcnt, gsum, ngroups, res};
        return launch_pipe<PGroupBy<kLdsGroups>>(p, cols, n, 0, L, s, "groupby");
    }
    PGroupBy<0>::Params p = {gcnt, gsum, ngroups, res};
    return launch_pipe<PGroupBy<0>>(p, cols, n, 0, L, s, "groupby");
}

}  // extern "C"
