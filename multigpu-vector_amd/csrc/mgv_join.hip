// mgv_join.hip -- hash-join entry points of the C ABI: build, probe (+ sum /
// digest), probe + group-by, probe + materialise.  The operator chains and the
// driver kernel are in mgv_pipes.hpp / mgv_ops.hpp.
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_pipes.hpp"

namespace mgv {

constexpr uint32_t kLdsGroupsSmall = 1024;
constexpr uint32_t kLdsGroups = 4096;

template <bool SEL, bool C, int G>
static int probe_group_g(const typename PProbeGroup<SEL, C, G>::Params &p, const Cols &cols, uint64_t n,
                         const mgv_launch &L, mgv_stream_t s) {
    return launch_pipe<PProbeGroup<SEL, C, G>>(p, cols, n, 0, L, s, "probe_groupby");
}

template <bool SEL, bool C>
static int probe_group_c(const uint64_t *slots, uint32_t nbuckets, int32_t bound, uint64_t *gcnt, int64_t *gsum,
                         uint32_t ngroups, const Cols &cols, uint64_t n, const mgv_launch &L, mgv_stream_t s) {
    if (ngroups <= kLdsGroupsSmall) {
        typename PProbeGroup<SEL, C, kLdsGroupsSmall>::Params p = {bound, slots, nbuckets, gcnt, gsum, ngroups};
        return probe_group_g<SEL, C, kLdsGroupsSmall>(p, cols, n, L, s);
    }
    if (ngroups <= kLdsGroups) {
        typename PProbeGroup<SEL, C, kLdsGroups>::Params p = {bound, slots, nbuckets, gcnt, gsum, ngroups};
        return probe_group_g<SEL, C, kLdsGroups>(p, cols, n, L, s);
    }
    typename PProbeGroup<SEL, C, 0>::Params p = {bound, slots, nbuckets, gcnt, gsum, ngroups};
    return probe_group_g<SEL, C, 0>(p, cols, n, L, s);
}

}  // namespace mgv

using namespace mgv;

#define MGV_REQUIRE_INIT()                                                            \
    do {                                                                              \
        if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called"); \
    } while (0)

extern "C" {

int mgv_hashjoin_build(uint64_t *slots, uint32_t nbuckets, const int32_t *keys, const int32_t *payload,
                       int payload_mode, const int32_t *pred, int32_t bound, uint64_t n, int64_t row_base,
                       mgv_result *res, mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!slots || !keys || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_hashjoin_build: null argument");
    if (n == 0) return MGV_OK;
    if (n >= (uint64_t)nbuckets * MGV_SLOTS_PER_BUCKET)
        return fail(MGV_E_CAPACITY, "mgv_hashjoin_build: %llu rows do not fit %u buckets", (unsigned long long)n,
                    nbuckets);
    const int32_t *pcol = nullptr;
    int rowid = -1;
    if (payload_mode == MGV_PAYLOAD_COLUMN) {
        if (!payload) return fail(MGV_E_INVALID, "mgv_hashjoin_build: payload column missing");
        pcol = payload;
    } else if (payload_mode == MGV_PAYLOAD_ROWID) {
        rowid = pred ? 2 : 1;
    } else if (payload_mode != MGV_PAYLOAD_NONE) {
        return fail(MGV_E_INVALID, "mgv_hashjoin_build: bad payload mode %d", payload_mode);
    }
    if (pred) {
        Cols cols = {{pred, keys, pcol, nullptr}, rowid};
        PBuild<true>::Params p = {bound, slots, nbuckets, res};
        return launch_pipe<PBuild<true>>(p, cols, n, row_base, L, s, "hashjoin_build");
    }
    Cols cols = {{keys, pcol, nullptr, nullptr}, rowid};
    PBuild<false>::Params p = {bound, slots, nbuckets, res};
    return launch_pipe<PBuild<false>>(p, cols, n, row_base, L, s, "hashjoin_build");
}

int mgv_hashjoin_probe(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                       const int32_t *keys, const int32_t *vals, uint64_t n, int64_t row_base, uint32_t flags,
                       mgv_result *res, mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!slots || !keys || !res || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_hashjoin_probe: null argument");
    if (n == 0) return MGV_OK;
    const bool digest = (flags & MGV_PROBE_DIGEST) != 0;
    const bool compact = L.compact != 0;  // default: compact before a probe
    if (pred) {
        Cols cols = {{pred, keys, vals, nullptr}, digest ? 3 : -1};
#define MGV_PS(C, D)                                                                      \
    {                                                                                     \
        typename PProbeSum<true, C, D>::Params p = {bound, slots, nbuckets, res};         \
        return launch_pipe<PProbeSum<true, C, D>>(p, cols, n, row_base, L, s, "probe"); \
    }
        if (compact && digest) MGV_PS(true, true)
        if (compact && !digest) MGV_PS(true, false)
        if (!compact && digest) MGV_PS(false, true)
        MGV_PS(false, false)
#undef MGV_PS
    }
    Cols cols = {{keys, vals, nullptr, nullptr}, digest ? 2 : -1};
    if (digest) {
        PProbeSum<false, true, true>::Params p = {bound, slots, nbuckets, res};
        return launch_pipe<PProbeSum<false, true, true>>(p, cols, n, row_base, L, s, "probe");
    }
    PProbeSum<false, true, false>::Params p = {bound, slots, nbuckets, res};
    if (L.reserved & MGV_LAUNCH_COOP_PROBE) {
        PProbeSum<false, true, false, OpProbeCoop>::Params pc = {bound, slots, nbuckets, res};
        return launch_pipe<PProbeSum<false, true, false, OpProbeCoop>>(pc, cols, n, row_base, L, s, "probe");
    }
    return launch_pipe<PProbeSum<false, true, false>>(p, cols, n, row_base, L, s, "probe");
}

int mgv_hashjoin_probe_groupby(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred,
                                          int32_t bound, const int32_t *keys, const int32_t *vals, uint64_t n,
                                          uint64_t *gcnt, int64_t *gsum, uint32_t ngroups, mgv_launch L,
                                          mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!slots || !keys || !vals || !gcnt || !gsum || nbuckets == 0 || ngroups == 0)
        return fail(MGV_E_INVALID, "mgv_hashjoin_probe_groupby: bad argument");
    if (n == 0) return MGV_OK;
    const bool compact = L.compact != 0;
    if (pred) {
        Cols cols = {{pred, keys, vals, nullptr}, -1};
        if (compact) return probe_group_c<true, true>(slots, nbuckets, bound, gcnt, gsum, ngroups, cols, n, L, s);
        return probe_group_c<true, false>(slots, nbuckets, bound, gcnt, gsum, ngroups, cols, n, L, s);
    }
    Cols cols = {{keys, vals, nullptr, nullptr}, -1};
    return probe_group_c<false, true>(slots, nbuckets, bound, gcnt, gsum, ngroups, cols, n, L, s);
}

int mgv_hashjoin_probe_materialize(const uint64_t *slots, uint32_t nbuckets, const int32_t *keys, uint64_t n,
                                   int64_t row_base, int32_t *out_r, int32_t *out_p, uint64_t cap, mgv_result *res,
                                   mgv_launch L, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!slots || !keys || !res || nbuckets == 0 || (cap && (!out_r || !out_p)))
        return fail(MGV_E_INVALID, "mgv_hashjoin_probe_materialize: bad argument");
    if (n == 0) return MGV_OK;
    Cols cols = {{keys, nullptr, nullptr, nullptr}, 1};
    PProbeMat::Params p = {slots, nbuckets, out_r, out_p, cap, res};
    return launch_pipe<PProbeMat>(p, cols, n, row_base, L, s, "probe_materialize");
}

}  // extern "C"
