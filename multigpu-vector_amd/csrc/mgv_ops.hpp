// mgv_ops.hpp -- device operators of the select -> hash-join -> aggregate path.
//
// Each operator keeps the lifecycle of the reference's device operators
// (consume_open / consume_warp / consume_close, src/operators/operator.cu:137-148):
//   consume_open()   block-wide, once per block before any row;
//   consume(Rows)    wave-converged, one call per wave chunk (<= 256 rows);
//   consume_close()  block-wide, once per block after the last row.
// Chains are composed statically (template parents) so the whole chain inlines
// into one kernel; the reference dispatched through a cross-TU `variant`
// visitor per 128-row chunk (operator.cu:137-148, Makefile:39 -rdc).
//
// Chunk layout.  A wave chunk is up to 4*64 rows; lane l holds rows 4l..4l+3
// (one dwordx4 per column), with a 4-bit validity mask.  The reference's warp
// chunk was 4*32 rows with row k*32+lane on lane `lane` (exchange.cu:185-188,
// select3.cu:44); the quad layout lets one global_load_dwordx4 per lane and
// column feed a chunk.
#pragma once

#include "mgv_wave.hpp"
#include "../../include/mgv.h"

namespace mgv {

constexpr uint64_t kEmptySlot = ~0ull;

// Lane-local slice of a wave chunk: 4 rows x up to 4 int32 columns.
struct Rows {
    v4i c[4];
    uint32_t m;  // bit j set: row j of this lane is valid
};

__device__ __forceinline__ int32_t q_get(const v4i &q, int j) {
    return j == 0 ? q.x : (j == 1 ? q.y : (j == 2 ? q.z : q.w));
}

// Bucket of a key: multiply-high range reduction of fmix32(key) onto
// [0, nbuckets).  With nbuckets = 2^k this selects the top k hash bits; the
// reference masks the low bits of the same hash (hashjoin.cu:52,146).
__host__ __device__ __forceinline__ uint32_t bucket_of(int32_t key, uint32_t nbuckets) {
    return (uint32_t)(((uint64_t)fmix32((uint32_t)key) * nbuckets) >> 32);
}

// ============================================================== aggregation

// SUM over the first NSUM columns (int64) and COUNT, optionally the join-pair
// digest over (c2 = probe row id, c0 = build payload).  Replaces
// aggregation<32,0> (aggregation.cu:19-66): the int32 LDS cell per thread
// becomes int64 registers, the __shfl_xor butterfly becomes DPP, and a block
// issues one atomic per result field instead of one per warp.
template <int NSUM, bool DIGEST, int WAVES>
struct OpAggregate {
    struct Smem {
        uint64_t part[4][WAVES];
    };
    Smem *sm;
    mgv_result *res;
    uint64_t cnt, sum, dsum, dxor;

    __device__ void consume_open() { cnt = sum = dsum = dxor = 0; }

    __device__ __forceinline__ void consume(const Rows &r) {
        cnt += (uint64_t)__popc(r.m & 0xFu);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((r.m >> j) & 1u) {
                int64_t s = 0;
#pragma unroll
                for (int i = 0; i < NSUM; ++i) s += (int64_t)q_get(r.c[i], j);
                sum += (uint64_t)s;
                if (DIGEST) {
                    uint64_t d = mix64(((uint64_t)(uint32_t)q_get(r.c[2], j) << 32) |
                                       (uint64_t)(uint32_t)q_get(r.c[0], j));
                    dsum += d;
                    dxor ^= d;
                }
            }
        }
    }

    __device__ void consume_close() {
        const uint32_t w = wave_id(), l = lane_id();
        uint64_t a = wave_sum_u64(cnt), b = wave_sum_u64(sum);
        uint64_t c = DIGEST ? wave_sum_u64(dsum) : 0, d = DIGEST ? wave_xor_u64(dxor) : 0;
        if (l == 0) {
            sm->part[0][w] = a;
            sm->part[1][w] = b;
            sm->part[2][w] = c;
            sm->part[3][w] = d;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t A = 0, B = 0, C = 0, D = 0;
            for (int i = 0; i < WAVES; ++i) {
                A += sm->part[0][i];
                B += sm->part[1][i];
                C += sm->part[2][i];
                D ^= sm->part[3][i];
            }
            if (A) {
                atomicAdd((unsigned long long *)&res->count, (unsigned long long)A);
                atomicAdd((unsigned long long *)&res->sum, (unsigned long long)B);
                if (DIGEST) {
                    atomicAdd((unsigned long long *)&res->digest_sum, (unsigned long long)C);
                    atomicXor((unsigned long long *)&res->digest_xor, (unsigned long long)D);
                }
            }
        }
        __syncthreads();
    }
};

// ============================================================== group-by

// GROUP BY c0 (group id) : COUNT, SUM(c1) -- no reference counterpart.  When
// LDS_GROUPS > 0 each block keeps a private table of LDS_GROUPS groups in LDS
// (ds_add_u32 / ds_add_u64) and flushes the non-empty groups with global
// atomics at close; otherwise rows go straight to global atomics.
template <int LDS_GROUPS>
struct OpGroupBy {
    struct Smem {
        uint32_t cnt[LDS_GROUPS > 0 ? LDS_GROUPS : 1];
        unsigned long long sum[LDS_GROUPS > 0 ? LDS_GROUPS : 1];
    };
    Smem *sm;
    uint64_t *gcnt;
    int64_t *gsum;
    uint32_t ngroups;
    mgv_result *res;
    uint64_t errs;

    __device__ void consume_open() {
        errs = 0;
        if (LDS_GROUPS > 0) {
            for (uint32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
                sm->cnt[g] = 0;
                sm->sum[g] = 0;
            }
            __syncthreads();
        }
    }

    __device__ __forceinline__ void consume(const Rows &r) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((r.m >> j) & 1u) {
                uint32_t g = (uint32_t)q_get(r.c[0], j);
                int64_t v = (int64_t)q_get(r.c[1], j);
                if (g < ngroups) {
                    if (LDS_GROUPS > 0) {
                        atomicAdd(&sm->cnt[g], 1u);
                        atomicAdd(&sm->sum[g], (unsigned long long)v);
                    } else {
                        atomicAdd((unsigned long long *)&gcnt[g], 1ull);
                        atomicAdd((unsigned long long *)&gsum[g], (unsigned long long)v);
                    }
                } else {
                    ++errs;
                }
            }
        }
    }

    __device__ void consume_close() {
        if (LDS_GROUPS > 0) {
            __syncthreads();
            for (uint32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
                uint32_t c = sm->cnt[g];
                if (c) {
                    atomicAdd((unsigned long long *)&gcnt[g], (unsigned long long)c);
                    atomicAdd((unsigned long long *)&gsum[g], sm->sum[g]);
                }
            }
        }
        uint64_t e = wave_sum_u64(errs);
        if (res && lane_id() == 0 && e) atomicAdd((unsigned long long *)&res->errors, (unsigned long long)e);
        __syncthreads();
    }
};

// ============================================================== materialise

// Appends column CR (and CP when TWO) of every valid row to out_a (/out_b)
// with one atomic reservation per wave chunk.  Replaces buffer::try_write /
// try_partial_final_write (src/buffer.cuh:92-134): the reservation covers the
// exact number of valid rows, so partial chunks need no rollback; rows past
// `cap` are counted in res->overflow and not written.
template <int CA, int CB, bool TWO>
struct OpMaterialize {
    struct Smem {
        int dummy;
    };
    Smem *sm;
    int32_t *out_a;
    int32_t *out_b;
    uint64_t cap;
    mgv_result *res;

    __device__ void consume_open() {}

    __device__ __forceinline__ void consume(const Rows &r) {
        uint32_t off[4];
        uint32_t tot = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint64_t b = ballot((r.m >> j) & 1u);
            off[j] = tot + mbcnt(b);
            tot += popc64(b);
        }
        if (tot == 0) return;
        unsigned long long base = 0;
        if (lane_id() == 0) {
            base = atomicAdd((unsigned long long *)&res->count, (unsigned long long)tot);
            unsigned long long end = base + tot;
            if (end > cap) {
                unsigned long long lost = end - (base > cap ? base : cap);
                atomicAdd((unsigned long long *)&res->overflow, lost);
            }
        }
        base = __builtin_amdgcn_readlane((uint32_t)base, 0) |
               ((unsigned long long)__builtin_amdgcn_readlane((uint32_t)(base >> 32), 0) << 32);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((r.m >> j) & 1u) {
                uint64_t pos = base + off[j];
                if (pos < cap) {
                    out_a[pos] = q_get(r.c[CA], j);
                    if (TWO) out_b[pos] = q_get(r.c[CB], j);
                }
            }
        }
    }

    __device__ void consume_close() {}
};

// ============================================================== hash join

// Probe: c0 = key, c1.. = carried columns.  For every slot whose key equals
// the probe key, the parent receives a row {c0 = slot payload, c1 = in.c1,
// c2 = in.c2}.  Replaces hashjoin::consume_warp (src/operators/hashjoin.cu:
// 135-182): the chain walk over first[]/next[] becomes a scan of 64-byte
// buckets, each read with four dwordx4 loads per lane (one cache line), for
// all four rows of a lane before any compare, and the per-lane atomicAdd of
// the count becomes the parent's register accumulation.
template <class Parent>
struct OpProbe {
    typedef typename Parent::Smem Smem;
    Parent parent;
    const uint64_t *slots;
    uint32_t nbuckets;

    __device__ void consume_open() { parent.consume_open(); }
    __device__ void consume_close() { parent.consume_close(); }

    __device__ __forceinline__ void consume(const Rows &in) {
        uint32_t b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = bucket_of(q_get(in.c[0], j), nbuckets);
        uint32_t pending = in.m & 0xFu;
        // Every pending row advances one bucket per pass, so nbuckets passes
        // cover the whole table even if it has no empty slot.
        for (uint32_t pass = 0; ballot(pending != 0) && pass < nbuckets; ++pass) {
            v4u q[4][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((pending >> j) & 1u) {
                    const v4u *p = reinterpret_cast<const v4u *>(slots + (uint64_t)b[j] * 8);
#pragma unroll
                    for (int t = 0; t < 4; ++t) q[j][t] = p[t];
                } else {
#pragma unroll
                    for (int t = 0; t < 4; ++t) q[j][t] = v4u{~0u, ~0u, ~0u, ~0u};
                }
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                Rows out;
                out.c[1] = in.c[1];
                out.c[2] = in.c[2];
                out.c[3] = in.c[3];
                uint32_t mm = 0;
                int32_t pay[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const v4u &w = q[j][s >> 1];
                    uint32_t k = (s & 1) ? w.z : w.x;
                    uint32_t p = (s & 1) ? w.w : w.y;
                    pay[j] = (int32_t)p;
                    bool hit = ((pending >> j) & 1u) && k == (uint32_t)q_get(in.c[0], j) && !(k == ~0u && p == ~0u);
                    mm |= (uint32_t)hit << j;
                }
                out.c[0] = v4i{pay[0], pay[1], pay[2], pay[3]};
                out.m = mm;
                if (ballot(mm != 0)) parent.consume(out);
            }
            // A bucket with an empty slot ends the scan (buckets fill as a
            // prefix); a full bucket continues into the next one.
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((pending >> j) & 1u) {
                    const v4u &w = q[j][3];
                    bool full = !(w.z == ~0u && w.w == ~0u);
                    if (full) {
                        b[j] = (b[j] + 1 == nbuckets) ? 0 : b[j] + 1;
                    } else {
                        pending &= ~(1u << j);
                    }
                }
            }
        }
    }
};

// Build: c0 = key, c1 = payload.  Replaces hashjoin_builder::consume_warp
// (src/operators/hashjoin.cu:41-60): instead of atomicExch on a chain head
// plus a separate {next,key} write, each row claims the first empty slot of
// its bucket run with one 64-bit CAS that writes key and payload together.
template <int WAVES>
struct OpBuild {
    struct Smem {
        int dummy;
    };
    Smem *sm;
    uint64_t *slots;
    uint32_t nbuckets;
    mgv_result *res;
    uint64_t errs;

    __device__ void consume_open() { errs = 0; }

    __device__ __forceinline__ void insert(uint32_t key, uint32_t pay) {
        const uint64_t val = ((uint64_t)pay << 32) | key;
        if (val == kEmptySlot) {
            ++errs;
            return;
        }
        uint32_t b = bucket_of((int32_t)key, nbuckets);
        // First guess: the first empty slot of the bucket as currently seen.
        const v4u *p = reinterpret_cast<const v4u *>(slots + (uint64_t)b * 8);
        uint32_t s = 8;
#pragma unroll
        for (int t = 3; t >= 0; --t) {
            v4u w = p[t];
            if (w.z == ~0u && w.w == ~0u) s = 2 * t + 1;
            if (w.x == ~0u && w.y == ~0u) s = 2 * t;
        }
        // At most one CAS per slot of the table, so a full table ends the loop.
        for (uint64_t tries = 0;; ++tries) {
            if (tries >= (uint64_t)nbuckets * 8) {
                ++errs;
                return;
            }
            if (s == 8) {
                b = (b + 1 == nbuckets) ? 0 : b + 1;
                s = 0;
            }
            unsigned long long old =
                atomicCAS((unsigned long long *)(slots + (uint64_t)b * 8 + s), (unsigned long long)kEmptySlot,
                          (unsigned long long)val);
            if (old == kEmptySlot) return;
            ++s;
        }
    }

    __device__ __forceinline__ void consume(const Rows &r) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if ((r.m >> j) & 1u) insert((uint32_t)q_get(r.c[0], j), (uint32_t)q_get(r.c[1], j));
    }

    __device__ void consume_close() {
        uint64_t e = wave_sum_u64(errs);
        if (res && lane_id() == 0 && e) atomicAdd((unsigned long long *)&res->errors, (unsigned long long)e);
    }
};

// ============================================================== select

// unstable_select with less_eq_than (src/operators/select3.cuh:10-48,
// select3.cu:19-270): rows with c0 <= bound pass on with their columns
// shifted down by one (c1.. -> c0..).
//
// COMPACT: wave64 stream compaction.  For each of the lane's four rows a
// 64-bit ballot and mbcnt give each survivor its slot in the wave's LDS stage
// (NOUT columns x 320 rows, the reference's 5*WARPSIZE); every 256 survivors
// the stage goes to the parent as one dense chunk, and the <= 63 rows beyond
// it move to the front.  consume_close hands each wave's remainder (< 256
// rows) to the parent as a partial chunk: the reference's block tree merge,
// global spill and last-block flush (select3.cu:87-231,255-270) only kept
// chunks dense for its 128-aligned buffer writes, which this library's
// parents do not need.
// !COMPACT: predication -- the parent gets the chunk with the survivors'
// validity bits; used when the parent is an aggregation.
template <class Parent, bool COMPACT, int NOUT, int WAVES>
struct OpSelect {
    struct Smem {
        int32_t stage[COMPACT ? WAVES : 1][COMPACT ? NOUT : 1][COMPACT ? 320 : 1];
        typename Parent::Smem parent;
    };
    Smem *sm;
    Parent parent;
    int32_t bound;
    uint32_t fill;  // wave-uniform count of staged rows

    __device__ void consume_open() {
        fill = 0;
        parent.consume_open();
    }

    __device__ __forceinline__ void emit_stage(uint32_t nrows) {
        const uint32_t w = wave_id(), l = lane_id();
        Rows out;
#pragma unroll
        for (int c = 0; c < 4; ++c) out.c[c] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < NOUT; ++c) out.c[c] = *reinterpret_cast<const v4i *>(&sm->stage[w][c][4 * l]);
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) m |= (uint32_t)(4 * l + j < nrows) << j;
        out.m = m;
        parent.consume(out);
    }

    __device__ __forceinline__ void consume(const Rows &in) {
        if (!COMPACT) {
            Rows out;
            out.c[0] = in.c[1];
            out.c[1] = in.c[2];
            out.c[2] = in.c[3];
            out.c[3] = v4i{0, 0, 0, 0};
            uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) m |= (uint32_t)(q_get(in.c[0], j) <= bound) << j;
            out.m = in.m & m;
            if (ballot(out.m != 0)) parent.consume(out);
            return;
        }
        const uint32_t w = wave_id(), l = lane_id();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bool pass = ((in.m >> j) & 1u) && q_get(in.c[0], j) <= bound;
            uint64_t bal = ballot(pass);
            if (pass) {
                uint32_t pos = fill + mbcnt(bal);
#pragma unroll
                for (int c = 0; c < NOUT; ++c) sm->stage[w][c][pos] = q_get(in.c[c + 1], j);
            }
            fill += popc64(bal);
            if (fill >= kChunk) {
                wave_sync();
                emit_stage(kChunk);
                wave_sync();
                uint32_t rest = fill - kChunk;
                int32_t mv[NOUT];
#pragma unroll
                for (int c = 0; c < NOUT; ++c) mv[c] = (l < rest) ? sm->stage[w][c][kChunk + l] : 0;
                wave_sync();
                if (l < rest) {
#pragma unroll
                    for (int c = 0; c < NOUT; ++c) sm->stage[w][c][l] = mv[c];
                }
                wave_sync();
                fill = rest;
            }
        }
    }

    __device__ void consume_close() {
        if (COMPACT && fill) {
            wave_sync();
            emit_stage(fill);
            fill = 0;
        }
        parent.consume_close();
    }
};

}  // namespace mgv
