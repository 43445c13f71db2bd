// mgv_probe_lds.hip -- hash-join probe against table slices held in LDS.
//
// The partitioned probe keeps a partition's table slice in the XCD's L2, but
// every probe row still costs one L1->L2 request, and those requests' latency
// and the L1's limit on outstanding misses bound it.  This path partitions
// the probe rows twice by the top bits of fmix32(key) -- the bits that also
// pick the bucket --:
//   level 1: F1 partitions (default 64), packed without gaps;
//   level 2: F2 = nbuckets / S partitions of the level-1 output, where a
//            slice of S <= 2048 buckets (128 KiB) fits in LDS; because the
//            level-1 output is already ordered by the top log2(F1) bits, a
//            level-2 tile spreads over only F2 / F1 partitions, so the second
//            scatter writes long runs.
// Then one block per CU claims level-2 partitions: it copies the partition's
// table slice (plus a few following buckets for chains that run past it)
// into LDS and probes the partition's rows from LDS; a chain that leaves the
// loaded range continues in global memory.  Semantics are those of
// mgv_hashjoin_probe: COUNT of matching (row, slot) pairs and SUM(val +
// payload), exact in 64 bits.  Replaces hashjoin::consume_warp
// (src/operators/hashjoin.cu:135-182) with its aggregation parent
// (src/operators/aggregation.cu:27-66).
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_ops.hpp"
#include "mgv_part.hpp"

namespace mgv {

constexpr uint32_t kSliceMax = 2048;     // buckets of a slice in LDS (128 KiB)
constexpr uint32_t kSliceExtra = 8;      // buckets loaded past the slice
constexpr int kLdsBlock = 1024;
// Quads (4 rows) each lane probes per step; the next step's quads are loaded
// while these are probed.  Measured on the 2^28-row probe: 2 per step ran
// within 1 % of 1 and 3 per step was 8 % slower (register spills), so the
// kernel is not bound by the row loads' latency.
constexpr int kAhead = 1;
constexpr uint64_t kQuadStride = 4ull * kLdsBlock;   // rows between a lane's quads

struct LdsProbeArgs {
    const uint64_t *slots;
    uint32_t nbuckets;
    uint32_t slice;          // buckets per level-2 partition
    PartLayout L;            // level-2 layout
    uint32_t *next;          // partition counter (zeroed before the launch)
    mgv_result *res;
};

// The probe rows reach the probe hashed: the first partitioning writes
// fmix32(key) in place of the key (the hash is computed there anyway), and
// fmix32 is a bijection, so a row hits a slot exactly when its hash equals
// the hash of the slot's key.  Slice keys are hashed once when a slice is
// loaded, and a row's bucket is one multiply of its hash.
__device__ __forceinline__ uint32_t bucket_of_hash(uint32_t h, uint32_t nb) {
    return (uint32_t)(((uint64_t)h * nb) >> 32);
}

// Scans one global-memory bucket (64 bytes: 8 {key, payload} slots, empty =
// (-1, -1)) for the key whose hash is h: number of hitting slots and the sum
// of their payloads.  The empty pattern is excluded from the compare, so key
// -1 is handled.
__device__ __forceinline__ void bucket_scan(const v4u (&q)[4], uint32_t h_key, uint32_t &hits, int64_t &psum) {
    uint32_t h = 0;
    int64_t t = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t sk = (s & 1) ? q[s >> 1].z : q[s >> 1].x;
        const uint32_t sp = (s & 1) ? q[s >> 1].w : q[s >> 1].y;
        const bool hit = fmix32(sk) == h_key && !(sk == ~0u && sp == ~0u);
        h += hit ? 1u : 0u;
        t += hit ? (int64_t)(int32_t)sp : 0;
    }
    psum += t;
    hits += h;
}

__device__ __forceinline__ bool bucket_full(const v4u (&q)[4]) { return !(q[3].z == ~0u && q[3].w == ~0u); }

// A table slice in LDS is stored keys first: per bucket its 8 keys (hashed,
// in two 16-byte quads), its 8 payloads, and its fill count c.  The build fills a
// bucket's slots strictly in order (a slot is claimed only after every lower
// slot was seen occupied, and slots are never freed), so the occupied slots
// are exactly [0, c): slot s hits key iff s < c and its key equals key (an
// empty slot never hits, whatever the key), and the bucket is full iff
// c == 8.  A probe row then reads 32 bytes of keys and 1 byte of count, and a
// payload only when it hits, instead of the 64-byte bucket.
struct Slice {
    v4u key[(kSliceMax + kSliceExtra) * 2];
    uint32_t pay[(kSliceMax + kSliceExtra) * 8];
    uint8_t cnt[kSliceMax + kSliceExtra];
};

// LDS quad of key quad w (slots 4w..4w+3) of slice bucket idx.  Buckets' key
// quads are 32 bytes apart, so quad w of every bucket would fall in one of 8
// bank groups; xor-ing w with bucket bit 3 spreads it over 16.
__device__ __forceinline__ uint32_t key_at(uint32_t idx, uint32_t w) { return idx * 2 + (w ^ ((idx >> 3) & 1u)); }

// Slot mask of the hits of key in a slice bucket with keys k0, k1 and fill
// count c.
__device__ __forceinline__ uint32_t slice_hits(const v4u &k0, const v4u &k1, uint32_t c, uint32_t key) {
    const uint32_t m = (uint32_t)(k0.x == key) | (uint32_t)(k0.y == key) << 1 | (uint32_t)(k0.z == key) << 2 |
                       (uint32_t)(k0.w == key) << 3 | (uint32_t)(k1.x == key) << 4 | (uint32_t)(k1.y == key) << 5 |
                       (uint32_t)(k1.z == key) << 6 | (uint32_t)(k1.w == key) << 7;
    return m & ((1u << c) - 1u);
}

// Payload sum of the hit slots hm of slice bucket idx, given the payload of
// the lowest hit slot (p1, read by the caller).  Duplicate build keys (two or
// more hits in one bucket) read the other payloads here.
__device__ __forceinline__ void slice_payloads(const Slice &S, uint32_t idx, uint32_t hm, int32_t p1, uint32_t &hits,
                                               int64_t &psum) {
    const uint32_t h = (uint32_t)__builtin_popcount(hm);
    int64_t t = h ? (int64_t)p1 : 0;
    if (ballot(h > 1)) {
        for (uint32_t m = hm & (hm - 1); m; m &= m - 1) t += (int64_t)(int32_t)S.pay[idx * 8 + __builtin_ctz(m)];
    }
    psum += t;
    hits += h;
}

__device__ __forceinline__ uint32_t lowest_slot(uint32_t hm) { return hm ? (uint32_t)__builtin_ctz(hm) : 0u; }

// Rows whose chain continues past their home bucket (about 5 % at load
// factor 0.5, but some row of nearly every wave) are not followed on the
// spot, which would cost the whole wave a scan per row: they are appended to
// a per-wave queue in LDS {key, value, bucket} and probed 64 at a time, one
// per lane, once the queue holds a full wave's worth (and at the end).
constexpr uint32_t kQueue = 128;   // entries per wave (a power of two)

struct WaveQueue {
    uint32_t key[kQueue];
    int32_t val[kQueue];
    uint32_t bkt[kQueue];
};

// Appends this lane's entry when `want`; returns the new tail.
__device__ __forceinline__ uint32_t queue_push(WaveQueue &Q, uint32_t tail, bool want, uint32_t key, int32_t val,
                                               uint32_t bkt) {
    const uint64_t bal = ballot(want);
    if (want) {
        const uint32_t pos = (tail + mbcnt(bal)) & (kQueue - 1);
        Q.key[pos] = key;
        Q.val[pos] = val;
        Q.bkt[pos] = bkt;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)(tail + (uint32_t)popc64(bal)));
}

// Probes up to 64 queued entries (one per lane) and re-queues the chains
// that go on; returns the new head and, through tail, the new tail.  A
// bucket inside the loaded range is read from LDS, any other from the table.
__device__ __forceinline__ uint32_t queue_round(WaveQueue &Q, uint32_t head, uint32_t &tail, const Slice &S,
                                                const uint64_t *slots, uint32_t nb, uint32_t lo, uint32_t nload,
                                                uint64_t &cnt, uint64_t &sum) {
    const uint32_t avail = tail - head;
    const uint32_t take = avail < kWave ? avail : kWave;
    const uint32_t l = lane_id();
    const bool mine = l < take;
    uint32_t key = 0, bkt = 0;
    int32_t val = 0;
    if (mine) {
        const uint32_t pos = (head + l) & (kQueue - 1);
        key = Q.key[pos];
        val = Q.val[pos];
        bkt = Q.bkt[pos];
    }
    wave_sync();   // every lane has read its entry before slots are reused
    const uint32_t idx = bkt - lo;
    uint32_t hits = 0;
    int64_t psum = 0;
    bool full = false;
    if (mine && idx < nload) {
        const v4u k0 = S.key[key_at(idx, 0)], k1 = S.key[key_at(idx, 1)];
        const uint32_t c = S.cnt[idx];
        const uint32_t hm = slice_hits(k0, k1, c, key);
        slice_payloads(S, idx, hm, (int32_t)S.pay[idx * 8 + lowest_slot(hm)], hits, psum);
        full = c == 8;
    } else if (mine) {
        const v4u *g = reinterpret_cast<const v4u *>(slots + (uint64_t)bkt * 8);
        v4u q[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) q[w] = load_global(g + w);
        bucket_scan(q, key, hits, psum);
        full = bucket_full(q);
    }
    if (mine && hits) {
        cnt += hits;
        sum += (uint64_t)(psum + (int64_t)hits * (int64_t)val);
    }
    // a chain ends at a bucket with an empty slot, or after visiting all nb
    // buckets once (when its next bucket is its home bucket again)
    const uint32_t nxt = (bkt + 1 == nb) ? 0 : bkt + 1;
    const bool more = mine && full && nxt != bucket_of_hash(key, nb);
    head += take;
    tail = queue_push(Q, tail, more, key, val, nxt);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)head);
}

// A lane's four rows: the four home buckets' keys and counts are read from
// LDS together (12 reads in flight) and compared, then the four lowest-hit
// payloads are read together; rows whose bucket is full continue at the next
// bucket, and rows whose home bucket lies outside the slice start at their
// home bucket, through the wave's queue.
__device__ __forceinline__ void lds_probe_quad(const v4i &k, const v4i &v, uint32_t m, const Slice &S,
                                               const uint64_t *slots, uint32_t nb, uint32_t lo, uint32_t nload,
                                               uint64_t &cnt, uint64_t &sum, WaveQueue &Q, uint32_t &head,
                                               uint32_t &tail) {
    uint32_t b[4], ii[4];
    bool ins[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        b[j] = bucket_of_hash((uint32_t)q_get(k, j), nb);
        const uint32_t idx = b[j] - lo;
        ins[j] = ((m >> j) & 1u) && idx < nload;
        ii[j] = ins[j] ? idx : 0u;
    }
    v4u k0[4], k1[4];
    uint32_t c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        k0[j] = S.key[key_at(ii[j], 0)];
        k1[j] = S.key[key_at(ii[j], 1)];
        c[j] = S.cnt[ii[j]];
    }
    uint32_t hm[4];
    int32_t p1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) hm[j] = ins[j] ? slice_hits(k0[j], k1[j], c[j], (uint32_t)q_get(k, j)) : 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) p1[j] = (int32_t)S.pay[ii[j] * 8 + lowest_slot(hm[j])];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t hits = 0;
        int64_t psum = 0;
        slice_payloads(S, ii[j], hm[j], p1[j], hits, psum);
        if (hits) {
            cnt += hits;
            sum += (uint64_t)(psum + (int64_t)hits * (int64_t)q_get(v, j));
        }
        const bool live = (m >> j) & 1u;
        const uint32_t from = ins[j] ? ((b[j] + 1 == nb) ? 0 : b[j] + 1) : b[j];
        const bool go = live && (!ins[j] || (c[j] == 8 && from != b[j]));
        tail = queue_push(Q, tail, go, (uint32_t)q_get(k, j), q_get(v, j), from);
        // fewer than kWave entries were queued before this push and at most
        // kWave were added, so the queue never holds more than kQueue - 1
        while (tail - head >= kWave) head = queue_round(Q, head, tail, S, slots, nb, lo, nload, cnt, sum);
    }
}

// Rows [i0, i0 + 4) of a partition ending at r1, as one unconditional
// dwordx4 load per column (no branch, so the loads of the next quad stay in
// flight while the current one is probed); the mask marks the rows < r1.
// Partition columns hold at least 4 rows past any in-range quad start
// (capacities are multiples of 64 rows), so the load stays in the buffer.
__device__ __forceinline__ uint32_t load_rows4(const int32_t *K, const int32_t *V, uint64_t i0, uint64_t r1, v4i &k,
                                               v4i &v) {
    k = load_stream4(K + i0);
    v = load_stream4(V + i0);
    const uint64_t left = r1 > i0 ? r1 - i0 : 0;
    return left >= 4 ? 0xfu : (1u << left) - 1u;
}

// Copies slice buckets [lo, lo + nload) of the table into S, one bucket per
// thread: keys, payloads and the fill count (the first empty slot, or 8).
__device__ __forceinline__ void load_slice(Slice &S, const uint64_t *slots, uint32_t lo, uint32_t nload) {
    for (uint32_t i = threadIdx.x; i < nload; i += kLdsBlock) {
        const v4u *g = reinterpret_cast<const v4u *>(slots + (uint64_t)(lo + i) * 8);
        const v4u q0 = load_global(g), q1 = load_global(g + 1), q2 = load_global(g + 2), q3 = load_global(g + 3);
        const v4u qs[4] = {q0, q1, q2, q3};
        uint32_t c = 8;
#pragma unroll
        for (int s = 7; s >= 0; --s) {
            const v4u &q = qs[s >> 1];
            const bool empty = (s & 1) ? (q.z == ~0u && q.w == ~0u) : (q.x == ~0u && q.y == ~0u);
            c = empty ? (uint32_t)s : c;
        }
        // hashed keys (an empty slot's value is never compared: it is >= c)
        S.key[key_at(i, 0)] = v4u{fmix32(q0.x), fmix32(q0.z), fmix32(q1.x), fmix32(q1.z)};
        S.key[key_at(i, 1)] = v4u{fmix32(q2.x), fmix32(q2.z), fmix32(q3.x), fmix32(q3.z)};
        v4u *pay4 = reinterpret_cast<v4u *>(S.pay);
        pay4[i * 2] = v4u{q0.y, q0.w, q1.y, q1.w};
        pay4[i * 2 + 1] = v4u{q2.y, q2.w, q3.y, q3.w};
        S.cnt[i] = (uint8_t)c;
    }
}

__global__ __launch_bounds__(kLdsBlock) void k_lds_probe(LdsProbeArgs a) {
    __shared__ __align__(16) Slice S;
    __shared__ WaveQueue s_queue[kLdsBlock / kWave];
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    WaveQueue &Q = s_queue[t / kWave];
    const uint32_t nb = a.nbuckets;
    const int32_t *K = a.L.col[0];
    const int32_t *V = a.L.col[1];
    uint64_t cnt = 0;
    uint64_t sum = 0;   // two's-complement int64
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t p = s_part;
        __syncthreads();   // s_part is rewritten next round; the slice is free again
        if (p >= a.L.F) break;
        const uint64_t r0 = a.L.pstart[p];
        const uint64_t r1 = r0 + a.L.ptotal[p];
        if (r0 == r1) continue;
        const uint32_t lo = p * a.slice;
        const uint32_t nload = min(a.slice + kSliceExtra, nb - lo);
        // the first rows' loads overlap the slice copy
        v4i k[kAhead], v[kAhead];
        uint32_t m[kAhead];
        uint64_t i0 = r0 + 4ull * t;
#pragma unroll
        for (int x = 0; x < kAhead; ++x) {
            const uint64_t ix = i0 + (uint64_t)x * kQuadStride;
            m[x] = load_rows4(K, V, ix < r1 ? ix : r0, ix < r1 ? r1 : r0, k[x], v[x]);
        }
        load_slice(S, a.slots, lo, nload);
        __syncthreads();
        // r0 is a multiple of 64 rows, so quads are 16-byte aligned
        uint32_t head = 0, tail = 0;   // the wave's queue (wave-uniform counters)
        // every lane of a wave iterates until the wave's last quad (lanes past
        // the end carry an empty mask), so the queue counters and ballots stay
        // wave-uniform
        while (ballot(i0 < r1)) {
            // the next step's quads load while this step's are probed
            const uint64_t i1 = i0 + (uint64_t)kAhead * kQuadStride;
            v4i kn[kAhead], vn[kAhead];
            uint32_t mn[kAhead];
#pragma unroll
            for (int x = 0; x < kAhead; ++x) {
                const uint64_t ix = i1 + (uint64_t)x * kQuadStride;
                mn[x] = load_rows4(K, V, ix < r1 ? ix : r0, ix < r1 ? r1 : r0, kn[x], vn[x]);
            }
#pragma unroll
            for (int x = 0; x < kAhead; ++x)
                lds_probe_quad(k[x], v[x], m[x], S, a.slots, nb, lo, nload, cnt, sum, Q, head, tail);
#pragma unroll
            for (int x = 0; x < kAhead; ++x) {
                k[x] = kn[x];
                v[x] = vn[x];
                m[x] = mn[x];
            }
            i0 = i1;
        }
        // the rest of the queue, before the slice is replaced
        while (tail != head) head = queue_round(Q, head, tail, S, a.slots, nb, lo, nload, cnt, sum);
        __syncthreads();
    }
    cnt = wave_sum_u64(cnt);
    sum = wave_sum_u64(sum);
    if (lane_id() == 0) {
        if (cnt) atomic_add_global((unsigned long long *)&a.res->count, (unsigned long long)cnt);
        if (sum) atomic_add_global((unsigned long long *)&a.res->sum, (unsigned long long)sum);
    }
}

// ------------------------------------------------------------------ build

struct LdsBuildArgs {
    uint64_t *slots;
    uint32_t nbuckets;
    uint32_t slice;          // buckets per level-2 partition
    PartLayout L;            // level-2 layout (key, payload)
    uint32_t *next;          // partition counter (zeroed before the launch)
    unsigned long long *ovf_count;
    int32_t *ovf_key;        // rows whose chain leaves their slice
    int32_t *ovf_pay;
    mgv_result *res;
};

// Inserts {key, pay} into slice bucket idx of an LDS slice of `slice`
// buckets, the way OpBuild::insert does in global memory: the first guess is
// the bucket's first empty slot as seen, a failed compare-and-swap moves to
// the next slot, a full bucket to the next bucket, so each bucket's slots are
// claimed in order.  Returns false when the chain would leave the slice (the
// caller finishes the row in the global table).
__device__ __forceinline__ bool lds_insert(unsigned long long *tab, uint32_t slice, uint32_t idx, uint64_t val) {
    const v4u *q = reinterpret_cast<const v4u *>(tab + (uint64_t)idx * 8);
    uint32_t s = 8;
#pragma unroll
    for (int t = 3; t >= 0; --t) {
        const v4u w = q[t];
        if (w.z == ~0u && w.w == ~0u) s = 2 * t + 1;
        if (w.x == ~0u && w.y == ~0u) s = 2 * t;
    }
    for (;;) {
        if (s == 8) {
            if (++idx == slice) return false;
            s = 0;
        }
        if (atomicCAS(&tab[idx * 8 + s], (unsigned long long)kEmptySlot, (unsigned long long)val) == kEmptySlot)
            return true;
        ++s;
    }
}

// One block per CU claims level-2 partitions: it copies the partition's
// table slice into LDS, inserts the partition's rows (all of which have
// their home bucket in the slice), and writes the slice back.  Rows whose
// chain runs past the slice go to the overflow list.
__global__ __launch_bounds__(kLdsBlock) void k_lds_build(LdsBuildArgs a) {
    __shared__ __align__(16) unsigned long long tab[kSliceMax * 8];
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    const int32_t *K = a.L.col[0];
    const int32_t *V = a.L.col[1];
    uint64_t errs = 0;
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t p = s_part;
        __syncthreads();
        if (p >= a.L.F) break;
        const uint64_t r0 = a.L.pstart[p];
        const uint64_t r1 = r0 + a.L.ptotal[p];
        if (r0 == r1) continue;   // an untouched slice is not copied
        const uint32_t lo = p * a.slice;
        v4u *g = reinterpret_cast<v4u *>(a.slots + (uint64_t)lo * 8);
        v4u *l = reinterpret_cast<v4u *>(tab);
        for (uint32_t i = t; i < a.slice * 4; i += kLdsBlock) l[i] = load_global(g + i);
        __syncthreads();
        // r0 is a multiple of 64 rows; columns are padded past the partition end
        for (uint64_t i0 = r0 + 4ull * t; i0 < r1; i0 += 4ull * kLdsBlock) {
            const v4i k = load_stream4(K + i0);
            const v4i v = load_stream4(V + i0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (i0 + j >= r1) break;
                const uint32_t key = (uint32_t)q_get(k, j), pay = (uint32_t)q_get(v, j);
                const uint64_t val = ((uint64_t)pay << 32) | key;
                if (val == kEmptySlot) {
                    ++errs;
                    continue;
                }
                // home buckets lie in the slice (bucket / slice = level-2
                // partition); anything else is finished in the global table
                const uint32_t idx = bucket_of((int32_t)key, a.nbuckets) - lo;
                if (idx >= a.slice || !lds_insert(tab, a.slice, idx, val)) {
                    const unsigned long long o = atomic_add_global(a.ovf_count, 1ull);
                    a.ovf_key[o] = (int32_t)key;
                    a.ovf_pay[o] = (int32_t)pay;
                }
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < a.slice * 4; i += kLdsBlock) g[i] = l[i];
        __syncthreads();
    }
    errs = wave_sum_u64(errs);
    if (lane_id() == 0 && errs && a.res) atomic_add_global((unsigned long long *)&a.res->errors, errs);
}

// The overflow rows, inserted into the global table after every slice is
// written back (their chains continue into the next slices, or wrap around
// to bucket 0).
__global__ __launch_bounds__(256) void k_lds_build_overflow(LdsBuildArgs a) {
    OpBuild<256 / kWave> op;
    op.slots = a.slots;
    op.nbuckets = a.nbuckets;
    op.res = a.res;
    op.consume_open();
    const uint64_t n = *a.ovf_count;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        op.insert((uint32_t)a.ovf_key[i], (uint32_t)a.ovf_pay[i]);
    op.consume_close();
}

// Level-2 fan-out for a table of nb buckets: the smallest power of two
// F2 >= 2 * F1 that divides nb into slices of at most kSliceMax buckets.
// When F1 alone already cuts the table into such slices (F1 > kMaxFanout
// asks for it), there is one level only: F2 = F1.
static int lds_fanout(uint32_t nb, uint32_t F1, uint32_t *F2) {
    if (F1 > kMaxFanout) {
        if (nb % F1 != 0 || nb / F1 > kSliceMax || nb / F1 == 0)
            return fail(MGV_E_INVALID, "LDS probe: one level of %u partitions does not cut %u buckets into slices "
                        "of <= %u buckets", F1, nb, kSliceMax);
        *F2 = F1;
        return MGV_OK;
    }
    uint32_t f = 2 * F1;
    while (f <= kMaxFanoutWide && (nb % f != 0 || nb / f > kSliceMax)) f <<= 1;
    if (f > kMaxFanoutWide || nb % f != 0 || nb / f > kSliceMax || nb / f == 0)
        return fail(MGV_E_INVALID, "LDS probe: %u buckets do not split into <= %u slices of <= %u buckets",
                    nb, kMaxFanoutWide, kSliceMax);
    *F2 = f;
    return MGV_OK;
}

// Workspace of the LDS paths: 256 bytes of counters, then the level-1 and
// level-2 partitionings (one level: only the second, at fan-out F1).
static int lds_plan(uint64_t n, uint32_t nb, uint32_t F1, void *ws, uint64_t ws_bytes, PartLayout *L1,
                    PartLayout *L2, uint32_t *F2, uint64_t *need) {
    if (F1 < 2 || F1 > kMaxFanoutWide || (F1 & (F1 - 1)))
        return fail(MGV_E_INVALID, "LDS probe: level-1 fan-out %u not a power of two in [2,%u]", F1, kMaxFanoutWide);
    int rc = lds_fanout(nb, F1, F2);
    if (rc) return rc;
    const bool one = *F2 == F1;
    uint64_t b1 = 0, b2 = 0;
    if (!one && (rc = plan_layout(n, F1, 2, nullptr, 0, nullptr, &b1, 1))) return rc;
    if ((rc = plan_layout(n, *F2, 2, nullptr, 0, nullptr, &b2, kPartAlign))) return rc;
    const uint64_t o1 = 256, o2 = 256 + (b1 + 255) / 256 * 256;
    if (need) *need = o2 + b2;
    if (!L1) return MGV_OK;
    if (!ws || ws_bytes < o2 + b2) return fail(MGV_E_CAPACITY, "LDS probe: workspace too small");
    if (!one && (rc = plan_layout(n, F1, 2, (char *)ws + o1, b1, L1, nullptr, 1))) return rc;
    if (one) L1->F = 0;   // marks the single level
    return plan_layout(n, *F2, 2, (char *)ws + o2, b2, L2, nullptr, kPartAlign);
}

// Build workspace: the probe's (counters + both partitionings of n rows)
// followed by the overflow list (n keys, n payloads).
static int build_plan(uint64_t n, uint32_t nb, uint32_t F1, void *ws, uint64_t ws_bytes, PartLayout *L1,
                      PartLayout *L2, uint32_t *F2, int32_t **okey, int32_t **opay, uint64_t *need) {
    uint64_t base = 0;
    int rc = lds_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, F2, &base);
    if (rc) return rc;
    base = (base + 255) / 256 * 256;
    const uint64_t ob = (n * 4 + 255) / 256 * 256;
    if (need) *need = base + 2 * ob;
    if (!L1) return MGV_OK;
    if (!ws || ws_bytes < base + 2 * ob) return fail(MGV_E_CAPACITY, "LDS build: workspace too small");
    if ((rc = lds_plan(n, nb, F1, ws, base, L1, L2, F2, nullptr))) return rc;
    *okey = (int32_t *)((char *)ws + base);
    *opay = (int32_t *)((char *)ws + base + ob);
    return MGV_OK;
}

// The probe of mgv_hashjoin_probe_lds on workspace ws (both entry points).
static int probe_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                     const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout, void *ws,
                     uint64_t ws_bytes, mgv_result *res, mgv_launch LC, mgv_stream_t st) {
    if (!slots || !keys || !res || nbuckets == 0) return fail(MGV_E_INVALID, "probe_lds: null argument");
    if (n == 0) return MGV_OK;
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "probe_lds: n must be < 2^31");
    PartLayout L1, L2;
    uint32_t F2 = 0;
    int rc = lds_plan(n, nbuckets, fanout, ws, ws_bytes, &L1, &L2, &F2, nullptr);
    if (rc) return rc;
    hipStream_t s = stream_of(st);
    // level 1: select + hash partitions, packed (the total is the survivor count)
    PartIn in1 = {pred, bound, keys, vals, 0, 0u, 0u, 0u, nullptr, nullptr, 1u};   // writes fmix32(key)
    if (L1.F == 0) {   // one level: straight into the slices' partitions
        if ((rc = partition_rows(L2, in1, s, scatter_shape(LC)))) return rc;
    } else {
        if ((rc = partition_rows(L1, in1, s, scatter_shape(LC)))) return rc;
        // level 2: the level-1 output, ordered by the same hash bits, split finer
        PartIn in2 = {nullptr, 0, L1.col[0], L1.col[1], 0, 0u, 0u, 0u, nullptr, L1.pstart + L1.F, 2u};   // hashes
        if ((rc = partition_rows(L2, in2, s, scatter_shape(LC)))) return rc;
    }
    uint32_t *next = (uint32_t *)ws;
    if ((rc = mgv_fill_u32(next, 1, 0, (mgv_stream_t)s))) return rc;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(k_lds_probe, kLdsBlock);
    const uint32_t grid = LC.grid ? LC.grid : (uint32_t)cached;
    LdsProbeArgs a = {slots, nbuckets, nbuckets / F2, L2, next, res};
    hipLaunchKernelGGL(k_lds_probe, dim3(grid), dim3(kLdsBlock), 0, s, a);
    return check_launch("k_lds_probe");
}

// Scratch of mgv_hashjoin_probe_lds_noalloc: the partition workspace of
// callers that cannot allocate.  The array lives in each GPU's memory for as
// long as the library is loaded; host code reaches it only through the
// device address hipGetSymbolAddress returns for the current device (kept in
// DeviceState::probe_scratch), never through the array's name.
constexpr uint64_t kProbeScratchBytes = MGV_PROBE_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kProbeScratchBytes];

int probe_scratch_setup(DeviceState &st) {
    void *p = nullptr;
    const int rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!p) return fail(MGV_E_NOT_INIT, "hipGetSymbolAddress(scratch) returned a null address");
    st.probe_scratch = p;
    return MGV_OK;
}

}  // namespace mgv

using namespace mgv;

extern "C" {

int mgv_probe_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes) {
    if (!bytes || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_probe_lds_workspace_size: bad argument");
    uint32_t F2 = 0;
    return lds_plan(n, nbuckets, fanout, nullptr, 0, nullptr, nullptr, &F2, bytes);
}

int mgv_hashjoin_probe_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                           const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout, void *ws,
                           uint64_t ws_bytes, mgv_result *res, mgv_launch LC, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    return probe_lds(slots, nbuckets, pred, bound, keys, vals, n, fanout, ws, ws_bytes, res, LC, st);
}

int mgv_hashjoin_probe_lds_noalloc(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout,
                                   mgv_result *res, mgv_launch LC, mgv_stream_t st) {
    DeviceState *ds = current_state();
    if (!ds) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!ds->probe_scratch) return fail(MGV_E_NOT_INIT, "probe_lds_noalloc: scratch array not resolved");
    if (!slots || !keys || !res || nbuckets == 0) return fail(MGV_E_INVALID, "probe_lds_noalloc: null argument");
    if (n == 0) return MGV_OK;
    uint32_t F2 = 0;
    uint64_t need = 0;
    int rc = lds_plan(n, nbuckets, fanout, nullptr, 0, nullptr, nullptr, &F2, &need);
    if (rc) return rc;
    if (need > kProbeScratchBytes)
        return fail(MGV_E_CAPACITY, "probe_lds_noalloc: %llu rows need %llu bytes of scratch, the array holds %llu",
                    (unsigned long long)n, (unsigned long long)need, (unsigned long long)kProbeScratchBytes);
    return probe_lds(slots, nbuckets, pred, bound, keys, vals, n, fanout, ds->probe_scratch, kProbeScratchBytes,
                     res, LC, st);
}

int mgv_build_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes) {
    if (!bytes || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_build_lds_workspace_size: bad argument");
    uint32_t F2 = 0;
    return build_plan(n, nbuckets, fanout, nullptr, 0, nullptr, nullptr, &F2, nullptr, nullptr, bytes);
}

int mgv_hashjoin_build_lds(uint64_t *slots, uint32_t nbuckets, const int32_t *keys, const int32_t *payload,
                           int payload_mode, const int32_t *pred, int32_t bound, uint64_t n, int64_t row_base,
                           uint32_t fanout, void *ws, uint64_t ws_bytes, mgv_result *res, mgv_launch LC,
                           mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!slots || !keys || nbuckets == 0) return fail(MGV_E_INVALID, "build_lds: null argument");
    if (n == 0) return MGV_OK;
    if (n >= (uint64_t)nbuckets * MGV_SLOTS_PER_BUCKET)
        return fail(MGV_E_CAPACITY, "build_lds: %llu rows do not fit %u buckets", (unsigned long long)n, nbuckets);
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "build_lds: n must be < 2^31");
    if (payload_mode == MGV_PAYLOAD_ROWID)   // row-id payloads: the direct build
        return mgv_hashjoin_build(slots, nbuckets, keys, payload, payload_mode, pred, bound, n, row_base, res, LC, st);
    if (payload_mode == MGV_PAYLOAD_COLUMN && !payload) return fail(MGV_E_INVALID, "build_lds: payload column missing");
    if (payload_mode != MGV_PAYLOAD_COLUMN && payload_mode != MGV_PAYLOAD_NONE)
        return fail(MGV_E_INVALID, "build_lds: bad payload mode %d", payload_mode);
    PartLayout L1, L2;
    uint32_t F2 = 0;
    int32_t *okey = nullptr, *opay = nullptr;
    int rc = build_plan(n, nbuckets, fanout, ws, ws_bytes, &L1, &L2, &F2, &okey, &opay, nullptr);
    if (rc) return rc;
    hipStream_t s = stream_of(st);
    const int32_t *pcol = payload_mode == MGV_PAYLOAD_COLUMN ? payload : nullptr;   // NONE: zeros
    PartIn in1 = {pred, bound, keys, pcol, 0, 0u, 0u, 0u, nullptr, nullptr};
    if (L1.F == 0) {
        if ((rc = partition_rows(L2, in1, s, scatter_shape(LC)))) return rc;
    } else {
        if ((rc = partition_rows(L1, in1, s, scatter_shape(LC)))) return rc;
        PartIn in2 = {nullptr, 0, L1.col[0], L1.col[1], 0, 0u, 0u, 0u, nullptr, L1.pstart + L1.F};
        if ((rc = partition_rows(L2, in2, s, scatter_shape(LC)))) return rc;
    }
    uint32_t *ctr = (uint32_t *)ws;   // [0]: partition counter, [2..3]: overflow count
    if ((rc = mgv_fill_u32(ctr, 4, 0, (mgv_stream_t)s))) return rc;
    LdsBuildArgs a = {slots, nbuckets, nbuckets / F2, L2, ctr, (unsigned long long *)(ctr + 2), okey, opay, res};
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(k_lds_build, kLdsBlock);
    hipLaunchKernelGGL(k_lds_build, dim3(LC.grid ? LC.grid : (uint32_t)cached), dim3(kLdsBlock), 0, s, a);
    if ((rc = check_launch("k_lds_build"))) return rc;
    hipLaunchKernelGGL(k_lds_build_overflow, dim3(256), dim3(256), 0, s, a);
    return check_launch("k_lds_build_overflow");
}

}  // extern "C"
