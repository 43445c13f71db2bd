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

// ------------------------------------------------------------------ launch

template <class Pipe, int B>
static int launch_b(const typename Pipe::Params &prm, const Cols &cols, uint64_t n, int64_t row_base,
                    const mgv_launch &L, hipStream_t s, const char *name) {
    auto kern = k_pipeline<Pipe, B>;
    uint64_t grid = L.grid;
    if (grid == 0) {
        static thread_local int cached = 0;
        if (!cached) cached = persistent_blocks(kern, B);
        grid = (uint64_t)cached;
        const uint64_t chunks_per_block = (uint64_t)(B / kWave);
        const uint64_t need = (n + (uint64_t)kChunk * chunks_per_block - 1) / ((uint64_t)kChunk * chunks_per_block);
        if (need < grid) grid = need ? need : 1;
    }
    if (grid > 0x7fffffffull) return fail(MGV_E_INVALID, "%s: grid too large", name);
    hipLaunchKernelGGL(kern, dim3((uint32_t)grid), dim3(B), 0, s, prm, cols, n, row_base);
    return check_launch(name);
}

template <class Pipe>
static int launch_pipe(const typename Pipe::Params &prm, const Cols &cols, uint64_t n, int64_t row_base,
                       const mgv_launch &L, mgv_stream_t st, const char *name) {
    hipStream_t s = stream_of(st);
    for (int c = 0; c < 4; ++c)
        if (cols.p[c] && !aligned16(cols.p[c]))
            return fail(MGV_E_ALIGN, "%s: column %d is not 16-byte aligned", name, c);
    switch (L.block) {
        case 64: return launch_b<Pipe, 64>(prm, cols, n, row_base, L, s, name);
        case 128: return launch_b<Pipe, 128>(prm, cols, n, row_base, L, s, name);
        case 0:
        case 256: return launch_b<Pipe, 256>(prm, cols, n, row_base, L, s, name);
        case 512: return launch_b<Pipe, 512>(prm, cols, n, row_base, L, s, name);
        case 1024: return launch_b<Pipe, 1024>(prm, cols, n, row_base, L, s, name);
        default: return fail(MGV_E_INVALID, "%s: block %u not in {64,128,256,512,1024}", name, L.block);
    }
}

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
    return launch_pipe<PProbeSum<false, true, false>>(p, cols, n, row_base, L, s, "probe");
}

}  // extern "C"

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

extern "C" int mgv_hashjoin_probe_groupby(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred,
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

extern "C" {

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
