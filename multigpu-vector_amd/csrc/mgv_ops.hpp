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

    __device__ __forceinline__ void consume_open() { cnt = sum = dsum = dxor = 0; }

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

    // One row per lane (probe output): c0 = payload, c1 = value, c2 = row id.
    static constexpr bool kUsesC2 = DIGEST;
    __device__ __forceinline__ void consume_one(int32_t c0, int32_t c1, int32_t c2, bool valid) {
        cnt += valid ? 1u : 0u;
        int64_t s = (int64_t)c0;
        if (NSUM > 1) s += (int64_t)c1;
        sum += valid ? (uint64_t)s : 0ull;
        if (DIGEST && valid) {
            uint64_t d = mix64(((uint64_t)(uint32_t)c2 << 32) | (uint64_t)(uint32_t)c0);
            dsum += d;
            dxor ^= d;
        }
    }

    __device__ __forceinline__ void consume_close() {
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
                atomic_add_global((unsigned long long *)&res->count, (unsigned long long)A);
                atomic_add_global((unsigned long long *)&res->sum, (unsigned long long)B);
                if (DIGEST) {
                    atomic_add_global((unsigned long long *)&res->digest_sum, (unsigned long long)C);
                    atomic_xor_global((unsigned long long *)&res->digest_xor, (unsigned long long)D);
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

    __device__ __forceinline__ void consume_open() {
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
                        atomic_add_lds(&sm->cnt[g], 1u);
                        atomic_add_lds(&sm->sum[g], (unsigned long long)v);
                    } else {
                        atomic_add_global((unsigned long long *)&gcnt[g], 1ull);
                        atomic_add_global((unsigned long long *)&gsum[g], (unsigned long long)v);
                    }
                } else {
                    ++errs;
                }
            }
        }
    }

    static constexpr bool kUsesC2 = false;
    __device__ __forceinline__ void consume_one(int32_t c0, int32_t c1, int32_t, bool valid) {
        const uint32_t g = (uint32_t)c0;
        if (valid) {
            if (g < ngroups) {
                if (LDS_GROUPS > 0) {
                    atomic_add_lds(&sm->cnt[g], 1u);
                    atomic_add_lds(&sm->sum[g], (unsigned long long)(int64_t)c1);
                } else {
                    atomic_add_global((unsigned long long *)&gcnt[g], 1ull);
                    atomic_add_global((unsigned long long *)&gsum[g], (unsigned long long)(int64_t)c1);
                }
            } else {
                ++errs;
            }
        }
    }

    __device__ __forceinline__ void consume_close() {
        if (LDS_GROUPS > 0) {
            __syncthreads();
            for (uint32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
                uint32_t c = sm->cnt[g];
                if (c) {
                    atomic_add_global((unsigned long long *)&gcnt[g], (unsigned long long)c);
                    atomic_add_global((unsigned long long *)&gsum[g], sm->sum[g]);
                }
            }
        }
        uint64_t e = wave_sum_u64(errs);
        if (res && lane_id() == 0 && e) atomic_add_global((unsigned long long *)&res->errors, (unsigned long long)e);
        __syncthreads();
    }
};

// Wave-level variants for the fused multi-query kernel, whose waves work on
// different queries at the same time: consume_close reduces within the wave
// (DPP) and issues one atomic per result field from lane 0, with no block
// barrier.
template <int NSUM, bool DIGEST>
struct OpAggWave {
    struct Smem {
        int dummy;
    };
    Smem *sm;
    mgv_result *res;
    uint64_t cnt, sum, dsum, dxor;

    __device__ __forceinline__ void consume_open() { cnt = sum = dsum = dxor = 0; }

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

    // One row per lane (probe output): c0 = payload, c1 = value, c2 = row id.
    static constexpr bool kUsesC2 = DIGEST;
    __device__ __forceinline__ void consume_one(int32_t c0, int32_t c1, int32_t c2, bool valid) {
        cnt += valid ? 1u : 0u;
        int64_t s = (int64_t)c0;
        if (NSUM > 1) s += (int64_t)c1;
        sum += valid ? (uint64_t)s : 0ull;
        if (DIGEST && valid) {
            uint64_t d = mix64(((uint64_t)(uint32_t)c2 << 32) | (uint64_t)(uint32_t)c0);
            dsum += d;
            dxor ^= d;
        }
    }

    __device__ __forceinline__ void consume_close() {
        const uint64_t a = wave_sum_u64(cnt), b = wave_sum_u64(sum);
        const uint64_t c = DIGEST ? wave_sum_u64(dsum) : 0, d = DIGEST ? wave_xor_u64(dxor) : 0;
        if (lane_id() == 0 && a) {
            atomic_add_global((unsigned long long *)&res->count, (unsigned long long)a);
            atomic_add_global((unsigned long long *)&res->sum, (unsigned long long)b);
            if (DIGEST) {
                atomic_add_global((unsigned long long *)&res->digest_sum, (unsigned long long)c);
                atomic_xor_global((unsigned long long *)&res->digest_xor, (unsigned long long)d);
            }
        }
    }
};

struct OpGroupWave {
    struct Smem {
        int dummy;
    };
    Smem *sm;
    uint64_t *gcnt;
    int64_t *gsum;
    uint32_t ngroups;
    mgv_result *res;
    uint64_t errs;

    __device__ __forceinline__ void consume_open() { errs = 0; }

    __device__ __forceinline__ void consume(const Rows &r) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((r.m >> j) & 1u) {
                const uint32_t g = (uint32_t)q_get(r.c[0], j);
                if (g < ngroups) {
                    atomic_add_global((unsigned long long *)&gcnt[g], 1ull);
                    atomic_add_global((unsigned long long *)&gsum[g], (unsigned long long)(int64_t)q_get(r.c[1], j));
                } else {
                    ++errs;
                }
            }
        }
    }

    static constexpr bool kUsesC2 = false;
    __device__ __forceinline__ void consume_one(int32_t c0, int32_t c1, int32_t, bool valid) {
        const uint32_t g = (uint32_t)c0;
        if (valid) {
            if (g < ngroups) {
                atomic_add_global((unsigned long long *)&gcnt[g], 1ull);
                atomic_add_global((unsigned long long *)&gsum[g], (unsigned long long)(int64_t)c1);
            } else {
                ++errs;
            }
        }
    }

    __device__ __forceinline__ void consume_close() {
        const uint64_t e = wave_sum_u64(errs);
        if (res && lane_id() == 0 && e) atomic_add_global((unsigned long long *)&res->errors, (unsigned long long)e);
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

    __device__ __forceinline__ void consume_open() {}

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
            base = atomic_add_global((unsigned long long *)&res->count, (unsigned long long)tot);
            unsigned long long end = base + tot;
            if (end > cap) {
                unsigned long long lost = end - (base > cap ? base : cap);
                atomic_add_global((unsigned long long *)&res->overflow, lost);
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

    static constexpr bool kUsesC2 = (CA == 2 || (TWO && CB == 2));
    __device__ __forceinline__ void consume_one(int32_t c0, int32_t c1, int32_t c2, bool valid) {
        const uint64_t b = ballot(valid);
        const uint32_t tot = popc64(b);
        if (tot == 0) return;
        unsigned long long base = 0;
        if (lane_id() == 0) {
            base = atomic_add_global((unsigned long long *)&res->count, (unsigned long long)tot);
            unsigned long long end = base + tot;
            if (end > cap) atomic_add_global((unsigned long long *)&res->overflow, end - (base > cap ? base : cap));
        }
        base = __builtin_amdgcn_readlane((uint32_t)base, 0) |
               ((unsigned long long)__builtin_amdgcn_readlane((uint32_t)(base >> 32), 0) << 32);
        const uint64_t pos = base + mbcnt(b);
        if (valid && pos < cap) {
            const int32_t v[3] = {c0, c1, c2};
            out_a[pos] = v[CA];
            if (TWO) out_b[pos] = v[CB];
        }
    }

    __device__ __forceinline__ void consume_close() {}
};

// ============================================================== hash join

// Probe, one row per lane: c0 = key, c1 = val, c2 = row id; for every slot
// whose key equals the probe key the parent receives one row {payload, val,
// row id} through consume_one.  Each lane hashes its four rows and reads
// their four 64-byte buckets with four dwordx4 loads each (16 loads in
// flight), compares the 8 keys of a bucket and selects the matching payload
// with a priority chain.  Per-row bookkeeping stays in scalar lane masks:
// hits, "two or more hits" (duplicate build keys, handled one slot at a
// time) and "bucket full" (continue into the next bucket) are built with
// s_or/s_and on the compare results.  Replaces hashjoin::consume_warp
// (src/operators/hashjoin.cu:135-182).
template <class Parent>
struct OpProbe {
    typedef typename Parent::Smem Smem;
    Parent parent;
    const uint64_t *slots;
    uint32_t nbuckets;

    __device__ __forceinline__ void consume_open() { parent.consume_open(); }
    __device__ __forceinline__ void consume_close() { parent.consume_close(); }

    // Scan one bucket (8 slots in q[0..3]) for `key`; emit the hits.
    // Returns whether the bucket is full (slot 7 used).  Rows that are not
    // pending carry an all-empty bucket and key 0, which matches nothing;
    // CAREFUL (some probe key is -1, the key half of an empty slot) adds the
    // payload half to the compare.
    template <bool CAREFUL>
    __device__ __forceinline__ bool scan(const v4u (&q)[4], uint32_t key, int32_t val, int32_t rid) {
        bool any = false, two = false;
        uint32_t sel = 0;
#pragma unroll
        for (int s = 7; s >= 0; --s) {
            const v4u &w = q[s >> 1];
            const uint32_t k = (s & 1) ? w.z : w.x;
            const uint32_t p = (s & 1) ? w.w : w.y;
            const bool h = CAREFUL ? (k == key && !(k == ~0u && p == ~0u)) : (k == key);
            two = two || (any && h);
            any = any || h;
            sel = h ? p : sel;   // lowest hitting slot wins
        }
        if (ballot(two)) {   // duplicate build keys: emit every hitting slot
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const v4u &w = q[s >> 1];
                const uint32_t k = (s & 1) ? w.z : w.x;
                const uint32_t p = (s & 1) ? w.w : w.y;
                const bool h = CAREFUL ? (k == key && !(k == ~0u && p == ~0u)) : (k == key);
                parent.consume_one((int32_t)p, val, rid, h);
            }
        } else {
            parent.consume_one((int32_t)sel, val, rid, any);
        }
        return !(q[3].z == ~0u && q[3].w == ~0u);
    }

    __device__ __forceinline__ void load_bucket(v4u (&q)[4], uint32_t b, bool pend) const {
        if (pend) {
            const v4u *p = reinterpret_cast<const v4u *>(slots + (uint64_t)b * 8);
#pragma unroll
            for (int t = 0; t < 4; ++t) q[t] = load_global(p + t);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) q[t] = v4u{~0u, ~0u, ~0u, ~0u};
        }
    }

    template <bool CAREFUL>
    __device__ __forceinline__ void consume_rows(const Rows &in, const uint32_t (&b)[4]) {
        uint32_t key[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) key[j] = ((in.m >> j) & 1u) ? (uint32_t)q_get(in.c[0], j) : 0u;
        v4u q[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) load_bucket(q[j], b[j], (in.m >> j) & 1u);
        uint32_t cont = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool full = scan<CAREFUL>(q[j], key[j], q_get(in.c[1], j), q_get(in.c[2], j));
            cont |= (uint32_t)(((in.m >> j) & 1u) && full) << j;
        }
        // rare: rows whose bucket was full continue into the next bucket
        if (ballot(cont != 0)) {
#pragma unroll 1
            for (int j = 0; j < 4; ++j) {
                bool pend = (cont >> j) & 1u;
                uint32_t bb = b[j];
                for (uint32_t pass = 1; ballot(pend) && pass < nbuckets; ++pass) {
                    bb = (bb + 1 == nbuckets) ? 0 : bb + 1;
                    v4u qq[4];
                    load_bucket(qq, bb, pend);
                    const bool full = scan<CAREFUL>(qq, pend ? key[j] : 0u, q_get(in.c[1], j), q_get(in.c[2], j));
                    pend = pend && full;
                }
            }
        }
    }

    __device__ __forceinline__ void consume(const Rows &in) {
        uint32_t b[4];
        bool neg = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int32_t k = q_get(in.c[0], j);
            b[j] = bucket_of(k, nbuckets);
            neg |= ((in.m >> j) & 1u) && k == -1;
        }
        if (ballot(neg)) consume_rows<true>(in, b);
        else consume_rows<false>(in, b);
    }
};

// Wave-cooperative probe (same contract as OpProbe).  Each lane hashes its
// own four rows; the chunk is then taken in 16 rounds in which the wave's 16
// groups of 4 lanes take one row each (key, bucket and value arrive by
// ds_bpermute) and every lane of a group reads one 16-byte quarter (2 slots)
// of the row's bucket with one dwordx4 load.  A load instruction thus touches
// 16 cache lines for 16 rows, where the one-row-per-lane probe touches 64
// lines for 16 rows' worth of buckets.  On MI355X this removes the
// texture-addresser and tag-conflict stalls of the partitioned probe (3.7x
// fewer L1 accesses) but not its time, which the L1->L2 request latency
// bounds; it is kept as a launch option (MGV_LAUNCH_COOP_PROBE).
template <class Parent>
struct OpProbeCoop {
    typedef typename Parent::Smem Smem;
    Parent parent;
    const uint64_t *slots;
    uint32_t nbuckets;

    __device__ __forceinline__ void consume_open() { parent.consume_open(); }
    __device__ __forceinline__ void consume_close() { parent.consume_close(); }

    __device__ __forceinline__ static int32_t bperm(uint32_t src_lane, int32_t v) {
        return __builtin_amdgcn_ds_bpermute((int)(src_lane << 2), v);
    }

    // Hits of this lane's two slots; returns "this group's bucket is full".
    template <bool CAREFUL>
    __device__ __forceinline__ bool round(const v4u &q, uint32_t key, int32_t val, int32_t rid, uint32_t grp,
                                          uint32_t sub) {
        const bool h0 = CAREFUL ? (q.x == key && !(q.x == ~0u && q.y == ~0u)) : (q.x == key);
        const bool h1 = CAREFUL ? (q.z == key && !(q.z == ~0u && q.w == ~0u)) : (q.z == key);
        if (ballot(h0 && h1)) {   // duplicate build keys in both slots
            parent.consume_one((int32_t)q.y, val, rid, h0);
            parent.consume_one((int32_t)q.w, val, rid, h1);
        } else {
            parent.consume_one((int32_t)(h0 ? q.y : q.w), val, rid, h0 || h1);
        }
        const uint64_t fb = ballot(sub == 3 && !(q.z == ~0u && q.w == ~0u));
        return (fb >> (4 * grp + 3)) & 1ull;
    }

    template <bool CAREFUL>
    __device__ __forceinline__ void consume_rows(const Rows &in, const int32_t (&bk)[4]) {
        const uint32_t l = lane_id();
        const uint32_t sub = l & 3u;
        const uint32_t grp = l >> 2;
        int32_t kin[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) kin[j] = ((in.m >> j) & 1u) ? q_get(in.c[0], j) : 0;
#pragma unroll 1
        for (int batch = 0; batch < 4; ++batch) {
            const uint32_t src = 16u * batch + grp;
            const uint32_t mv = (uint32_t)bperm(src, (int32_t)in.m) & 0xFu;
            if (!ballot(mv != 0)) continue;
            uint32_t key[4], b[4];
            int32_t val[4], rid[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                key[j] = (uint32_t)bperm(src, kin[j]);
                b[j] = (uint32_t)bperm(src, bk[j]);
                val[j] = bperm(src, q_get(in.c[1], j));
                rid[j] = Parent::kUsesC2 ? bperm(src, q_get(in.c[2], j)) : 0;
            }
            v4u q[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((mv >> j) & 1u) q[j] = load_global(reinterpret_cast<const v4u *>(slots + (uint64_t)b[j] * 8 + 2 * sub));
                else q[j] = v4u{~0u, ~0u, ~0u, ~0u};
            }
            uint32_t cont = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                cont |= (uint32_t)(round<CAREFUL>(q[j], key[j], val[j], rid[j], grp, sub) && ((mv >> j) & 1u)) << j;
            if (ballot(cont != 0)) {
#pragma unroll 1
                for (int j = 0; j < 4; ++j) {
                    bool pend = (cont >> j) & 1u;
                    uint32_t bb = b[j];
                    for (uint32_t pass = 1; ballot(pend) && pass < nbuckets; ++pass) {
                        bb = (bb + 1 == nbuckets) ? 0 : bb + 1;
                        v4u qq = v4u{~0u, ~0u, ~0u, ~0u};
                        if (pend) qq = load_global(reinterpret_cast<const v4u *>(slots + (uint64_t)bb * 8 + 2 * sub));
                        pend = round<CAREFUL>(qq, pend ? key[j] : 0u, val[j], rid[j], grp, sub) && pend;
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void consume(const Rows &in) {
        int32_t bk[4];
        bool neg = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int32_t k = q_get(in.c[0], j);
            bk[j] = (int32_t)bucket_of(k, nbuckets);
            neg |= ((in.m >> j) & 1u) && k == -1;
        }
        if (ballot(neg)) consume_rows<true>(in, bk);
        else consume_rows<false>(in, bk);
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

    __device__ __forceinline__ void consume_open() { errs = 0; }

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
            v4u w = load_global(p + t);
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
                atomic_cas_global((unsigned long long *)(slots + (uint64_t)b * 8 + s), (unsigned long long)kEmptySlot,
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

    __device__ __forceinline__ void consume_close() {
        uint64_t e = wave_sum_u64(errs);
        if (res && lane_id() == 0 && e) atomic_add_global((unsigned long long *)&res->errors, (unsigned long long)e);
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

    __device__ __forceinline__ void consume_open() {
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

    __device__ __forceinline__ void consume_close() {
        if (COMPACT && fill) {
            wave_sync();
            emit_stage(fill);
            fill = 0;
        }
        parent.consume_close();
    }
};

}  // namespace mgv
