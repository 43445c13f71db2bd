// mgv_pipes.hpp -- pipeline driver and operator chains shared by the
// per-vector / resident launches and the partitioned probe.
//
// k_pipeline<Pipe, BLOCK> is the MI355X counterpart of launch_consume_pipeline
// (src/operators/exchange.cu:124-194): every block runs consume_open, its
// waves walk the rows in 256-row wave chunks with a grid stride, then
// consume_close.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mgv_internal.hpp"
#include "mgv_ops.hpp"

namespace mgv {

// Input columns of a pipeline: up to 4 int32 columns; column `rowid_col`
// (if >= 0) is synthesised as row_base + row index instead of being loaded.
struct Cols {
    const int32_t *p[4];
    int32_t rowid_col;
};

__device__ __forceinline__ void load_rows(Rows &r, const Cols &cols, uint64_t base, uint64_t n, int64_t row_base) {
    const uint64_t i0 = base + 4ull * lane_id();
    if (i0 + 3 < n) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (cols.p[c]) {
                r.c[c] = load_stream4(cols.p[c] + i0);
            } else if (c == cols.rowid_col) {
                int32_t b = (int32_t)(row_base + (int64_t)i0);
                r.c[c] = v4i{b, b + 1, b + 2, b + 3};
            } else {
                r.c[c] = v4i{0, 0, 0, 0};
            }
        }
        r.m = 0xFu;
    } else {
        int32_t v[4][4];
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t i = i0 + j;
            const bool ok = i < n;
            m |= (uint32_t)ok << j;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                int32_t x = 0;
                if (ok) {
                    if (cols.p[c]) x = load_stream1(cols.p[c] + i);
                    else if (c == cols.rowid_col) x = (int32_t)(row_base + (int64_t)i);
                }
                v[c][j] = x;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) r.c[c] = v4i{v[c][0], v[c][1], v[c][2], v[c][3]};
        r.m = m;
    }
}

// Whole chunks (base + kChunk <= n) of a pipeline that reads columns
// [0, NC) are loaded in two steps.  load_full issues one dwordx4 per lane and
// column, unconditionally (a null column loads column 0 instead); finish_full,
// called only when the chunk is about to be consumed, zeroes the null
// columns and fills in row ids.  Nothing touches the loaded registers between
// the two, so the next chunk's loads stay in flight while the current chunk
// is consumed (a branch or a select right after the loads makes the compiler
// wait for them at once).  Columns from NC on are row ids or zeros.
template <int NC>
__device__ __forceinline__ void load_full(Rows &r, const Cols &cols, uint64_t base) {
    const uint64_t i0 = base + 4ull * lane_id();
#pragma unroll
    for (int c = 0; c < NC; ++c) r.c[c] = load_stream4((cols.p[c] ? cols.p[c] : cols.p[0]) + i0);
}

template <int NC>
__device__ __forceinline__ void finish_full(Rows &r, const Cols &cols, uint64_t base, int64_t row_base) {
    const int32_t b = (int32_t)(row_base + (int64_t)(base + 4ull * lane_id()));
    const v4i ids = v4i{b, b + 1, b + 2, b + 3};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (c < NC) {
            const int32_t keep = cols.p[c] ? -1 : 0;
            const v4i x = r.c[c];
            r.c[c] = (c == cols.rowid_col) ? ids : v4i{x.x & keep, x.y & keep, x.z & keep, x.w & keep};
        } else {
            r.c[c] = (c == cols.rowid_col) ? ids : v4i{0, 0, 0, 0};
        }
    }
    r.m = 0xFu;
}

// Marks a whole chunk's loaded registers as read at this point.  Called right
// after the chunk is consumed: a consumer that skips a column (no lane passed
// the filter) would otherwise leave its load pending until the registers are
// next written, and the wait there would also cover the loads issued since.
template <int NC>
__device__ __forceinline__ void settle(const Rows &r) {
#pragma unroll
    for (int c = 0; c < NC; ++c) asm volatile("" ::"v"(r.c[c].x), "v"(r.c[c].y), "v"(r.c[c].z), "v"(r.c[c].w));
}

// Walks the wave's chunks first, first + stride, ... below n (first, stride
// and n wave-uniform, stride >= kChunk), consuming chunk i while chunk i+1 is
// loaded.  Two row buffers alternate, so a buffer is never copied while its
// loads are in flight; the loop handles a pair of whole chunks per iteration
// and has a single exit, and its prefetches are unconditional (the last one
// of an even count reloads the current chunk, in bounds, result unused): a
// conditional load or a second loop exit makes the compiler wait for every
// load before each consume.  Only the last chunk can be partial; it uses the
// bounds-checked load_rows.
template <int NC, class Op>
__device__ __forceinline__ void drive_chunks(Op &op, const Cols &cols, uint64_t first, uint64_t stride, uint64_t n,
                                             int64_t row_base) {
    const uint64_t whole = first + kChunk <= n ? (n - kChunk - first) / stride + 1 : 0;
    uint64_t base = first;
    if (whole) {
        Rows A, B;
        load_full<NC>(A, cols, base);
        for (uint64_t k = 2; k <= whole; k += 2) {
            const uint64_t b1 = base + stride, b2 = b1 + stride;
            load_full<NC>(B, cols, b1);
            finish_full<NC>(A, cols, base, row_base);
            op.consume(A);
            settle<NC>(A);
            load_full<NC>(A, cols, k < whole ? b2 : b1);
            finish_full<NC>(B, cols, b1, row_base);
            op.consume(B);
            settle<NC>(B);
            base = b2;
        }
        if (whole & 1) {
            finish_full<NC>(A, cols, base, row_base);
            op.consume(A);
        }
        base = first + whole * stride;
    }
    if (base < n) {
        Rows T;
        load_rows(T, cols, base, n, row_base);
        op.consume(T);
    }
}

// Each wave walks its chunks with a grid stride, loading chunk i+1 while it
// consumes chunk i.
template <class Pipe, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pipeline(typename Pipe::Params prm, Cols cols, uint64_t n,
                                                    int64_t row_base) {
    constexpr int WAVES = BLOCK / kWave;
    typedef typename Pipe::template Op<WAVES> Op;
    __shared__ __align__(16) typename Op::Smem sm;
    Op op;
    Pipe::setup(op, prm, &sm);
    op.consume_open();
    const uint64_t w = (uint64_t)blockIdx.x * WAVES + wave_id();
    const uint64_t stride = (uint64_t)gridDim.x * WAVES * kChunk;
    // The next chunk's loads are issued before the current chunk is consumed.
    // With a grid stride this one-ahead loop (whose consume also waits for
    // the chunk just issued) measured as fast as drive_chunks with 1024-thread
    // blocks and 9 % faster with 256-thread blocks on a 2^28-row scan; the
    // deeper overlap of drive_chunks pays off on the contiguous slices of the
    // fused and partitioned kernels instead.
    uint64_t base = w * kChunk;
    Rows cur;
    if (base < n) load_rows(cur, cols, base, n, row_base);
    for (; base < n; base += stride) {
        Rows nxt;
        const uint64_t nb = base + stride;
        if (nb < n) load_rows(nxt, cols, nb, n, row_base);
        op.consume(cur);
        cur = nxt;
    }
    op.consume_close();
}

// ------------------------------------------------------------------ pipelines

// SELECT COUNT(*), SUM(c1) WHERE c0 <= bound.
template <bool COMPACT>
struct PSelectSum {
    static constexpr int kCols = 2;   // pred, val
    struct Params {
        int32_t bound;
        mgv_result *res;
    };
    template <int W>
    using Op = OpSelect<OpAggregate<1, false, W>, COMPACT, 1, W>;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        op.sm = sm;
        op.bound = p.bound;
        op.parent.sm = &sm->parent;
        op.parent.res = p.res;
    }
};

// SELECT c1 WHERE c0 <= bound, materialised.
struct PSelectMat {
    static constexpr int kCols = 2;   // pred, val
    struct Params {
        int32_t bound;
        int32_t *out;
        uint64_t cap;
        mgv_result *res;
    };
    template <int W>
    using Op = OpSelect<OpMaterialize<0, 0, false>, true, 1, W>;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        op.sm = sm;
        op.bound = p.bound;
        op.parent.sm = &sm->parent;
        op.parent.out_a = p.out;
        op.parent.out_b = nullptr;
        op.parent.cap = p.cap;
        op.parent.res = p.res;
    }
};

// Build: (c0 key, c1 payload) or, with SEL, (c0 pred, c1 key, c2 payload).
template <bool SEL>
struct PBuild {
    static constexpr int kCols = SEL ? 3 : 2;   // [pred], key, payload
    struct Params {
        int32_t bound;
        uint64_t *slots;
        uint32_t nbuckets;
        mgv_result *res;
    };
    template <int W>
    using Op = typename std::conditional<SEL, OpSelect<OpBuild<W>, true, 2, W>, OpBuild<W>>::type;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        if constexpr (SEL) {
            op.sm = sm;
            op.bound = p.bound;
            op.parent.sm = &sm->parent;
            op.parent.slots = p.slots;
            op.parent.nbuckets = p.nbuckets;
            op.parent.res = p.res;
        } else {
            op.sm = sm;
            op.slots = p.slots;
            op.nbuckets = p.nbuckets;
            op.res = p.res;
        }
    }
};

// Probe + COUNT/SUM(val + payload) [+ digest]: (c0 key, c1 val, c2 rowid) or,
// with SEL, (c0 pred, c1 key, c2 val, c3 rowid).
template <bool SEL, bool COMPACT, bool DIGEST, template <class> class PROBE = OpProbe>
struct PProbeSum {
    static constexpr int kCols = (SEL ? 3 : 2) + (DIGEST ? 1 : 0);   // [pred], key, val, [row id]
    struct Params {
        int32_t bound;
        const uint64_t *slots;
        uint32_t nbuckets;
        mgv_result *res;
    };
    template <int W>
    using Agg = OpAggregate<2, DIGEST, W>;
    template <int W>
    using Op = typename std::conditional<SEL, OpSelect<PROBE<Agg<W>>, COMPACT, DIGEST ? 3 : 2, W>,
                                         PROBE<Agg<W>>>::type;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        if constexpr (SEL) {
            op.sm = sm;
            op.bound = p.bound;
            op.parent.slots = p.slots;
            op.parent.nbuckets = p.nbuckets;
            op.parent.parent.sm = &sm->parent;
            op.parent.parent.res = p.res;
        } else {
            op.slots = p.slots;
            op.nbuckets = p.nbuckets;
            op.parent.sm = sm;
            op.parent.res = p.res;
        }
    }
};

// Probe + GROUP BY payload: (c0 key, c1 val) or (c0 pred, c1 key, c2 val).
template <bool SEL, bool COMPACT, int LDSG>
struct PProbeGroup {
    static constexpr int kCols = SEL ? 3 : 2;   // [pred], key, val
    struct Params {
        int32_t bound;
        const uint64_t *slots;
        uint32_t nbuckets;
        uint64_t *gcnt;
        int64_t *gsum;
        uint32_t ngroups;
    };
    template <int W>
    using G = OpGroupBy<LDSG>;
    template <int W>
    using Op = typename std::conditional<SEL, OpSelect<OpProbe<G<W>>, COMPACT, 2, W>, OpProbe<G<W>>>::type;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        if constexpr (SEL) {
            op.sm = sm;
            op.bound = p.bound;
            op.parent.slots = p.slots;
            op.parent.nbuckets = p.nbuckets;
            auto &g = op.parent.parent;
            g.sm = &sm->parent;
            g.gcnt = p.gcnt;
            g.gsum = p.gsum;
            g.ngroups = p.ngroups;
            g.res = nullptr;
        } else {
            op.slots = p.slots;
            op.nbuckets = p.nbuckets;
            auto &g = op.parent;
            g.sm = sm;
            g.gcnt = p.gcnt;
            g.gsum = p.gsum;
            g.ngroups = p.ngroups;
            g.res = nullptr;
        }
    }
};

// Probe + materialise (r_row, payload): (c0 key, c1 rowid).
struct PProbeMat {
    static constexpr int kCols = 1;   // key (row ids computed)
    struct Params {
        const uint64_t *slots;
        uint32_t nbuckets;
        int32_t *out_r;
        int32_t *out_p;
        uint64_t cap;
        mgv_result *res;
    };
    template <int W>
    using Op = OpProbe<OpMaterialize<1, 0, true>>;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        op.slots = p.slots;
        op.nbuckets = p.nbuckets;
        op.parent.sm = sm;
        op.parent.out_a = p.out_r;
        op.parent.out_b = p.out_p;
        op.parent.cap = p.cap;
        op.parent.res = p.res;
    }
};

// GROUP BY c0: COUNT, SUM(c1).
template <int LDSG>
struct PGroupBy {
    static constexpr int kCols = 2;   // key, val
    struct Params {
        uint64_t *gcnt;
        int64_t *gsum;
        uint32_t ngroups;
        mgv_result *res;
    };
    template <int W>
    using Op = OpGroupBy<LDSG>;
    template <class O>
    __device__ static void setup(O &op, const Params &p, typename O::Smem *sm) {
        op.sm = sm;
        op.gcnt = p.gcnt;
        op.gsum = p.gsum;
        op.ngroups = p.ngroups;
        op.res = p.res;
    }
};

// ------------------------------------------------------------------ launch

template <class Pipe, int B>
inline int launch_b(const typename Pipe::Params &prm, const Cols &cols, uint64_t n, int64_t row_base,
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
inline int launch_pipe(const typename Pipe::Params &prm, const Cols &cols, uint64_t n, int64_t row_base,
                       const mgv_launch &L, mgv_stream_t st, const char *name) {
    hipStream_t s = stream_of(st);
    for (int c = 0; c < 4; ++c)
        if (cols.p[c] && !aligned16(cols.p[c]))
            return fail(MGV_E_ALIGN, "%s: column %d is not 16-byte aligned", name, c);
    switch (L.block) {
        case 64: return launch_b<Pipe, 64>(prm, cols, n, row_base, L, s, name);
        case 0:
        case 256: return launch_b<Pipe, 256>(prm, cols, n, row_base, L, s, name);
        case 1024: return launch_b<Pipe, 1024>(prm, cols, n, row_base, L, s, name);
        default: return fail(MGV_E_INVALID, "%s: block %u not in {64,256,1024}", name, L.block);
    }
}

}  // namespace mgv
