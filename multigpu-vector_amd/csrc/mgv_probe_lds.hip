// mgv_probe_lds.hip -- hash-join probe against table slices held in LDS.
//
// The partitioned probe keeps a partition's table slice in the XCD's L2, but
// every probe row still costs one L1->L2 request, and those requests' latency
// and the L1's limit on outstanding misses bound it.  This path partitions
// the probe rows by the top bits of fmix32(key) -- the bits that also pick
// the bucket -- into F2 = nbuckets / S partitions, where a slice of S <= 2048
// buckets fits in LDS:
//   level 1: F1 regions of fixed capacity, filled without count passes (one
//            atomic reservation per tile and partition; rows past a full
//            region go to an overflow area);
//   level 2: each level-1 region cut into F2 / F1 regions the same way (a
//            level-1 region holds one value of the top log2(F1) hash bits,
//            so a level-2 tile spans only F2 / F1 partitions and its runs
//            stay long); one level when nbuckets / F1 already is a slice.
// Then one block per CU claims level-2 partitions: it copies the partition's
// table slice (plus a few following buckets for chains that run past it)
// into LDS -- as 16-bit slot tags for power-of-two tables of >= 2^18
// buckets, else as key hashes -- and probes the partition's rows from LDS; a
// chain that leaves the loaded range continues in global memory, and the
// overflow areas' rows are probed from the table directly.  Semantics are
// those of mgv_hashjoin_probe: COUNT of matching (row, slot) pairs and
// SUM(val + payload), exact in 64 bits.  Replaces hashjoin::consume_warp
// (src/operators/hashjoin.cu:135-182) with its aggregation parent
// (src/operators/aggregation.cu:27-66).
#include <hip/hip_runtime.h>

#include "mgv_batch.hpp"
#include "mgv_internal.hpp"
#include "mgv_ops.hpp"
#include "mgv_part.hpp"

namespace mgv {

constexpr uint32_t kSliceMax = 2048;     // buckets of a slice in LDS (128 KiB)
constexpr uint32_t kSliceExtra = 8;      // buckets loaded past the slice
constexpr uint32_t kProbeSlice = kSliceMax;   // slice of mgv_hashjoin_probe_lds (count-free)
constexpr int kLdsBlock = 1024;
// Quads (4 rows) each lane probes per step; the next step's quads are loaded
// while these are probed.  Measured on the 2^28-row probe: 2 per step ran
// within 1 % of 1 and 3 per step was 8 % slower (register spills), so the
// kernel is not bound by the row loads' latency.
constexpr int kAhead = 1;
constexpr uint64_t kQuadStride = 4ull * kLdsBlock;   // rows between a lane's quads

// Tagged slices (tables of nb = 2^B buckets, B >= 18): a bucket's home is
// the top B bits of a key's hash, so a slot of bucket b holding a key of
// hash hk is identified exactly by its chain offset j = b - (hk >> L)
// (L = 32 - B) and the low L bits of hk.  With J = B - 16 bits for j, the
// tag (j << L) | (hk & (2^L - 1)) fits 16 bits while j <= 2^J - 2; a row of
// hash h at bucket b computes its own tag with j = b - (h >> L), and a slot
// hits exactly when the tags are equal.  No row computes a tag whose j bits
// are all ones: 0xFFFF marks an empty slot and 0xFFFE a slot whose j does
// not fit (a row reaches it only at j > 2^J - 2 and continues in global
// memory).
struct TagGeom {
    uint32_t L;         // low hash bits in a tag
    uint32_t lowmask;   // 2^L - 1
    uint32_t jmax;      // largest chain offset a tag holds: 2^(16 - L) - 2
};

constexpr uint32_t kTagEmpty = 0xFFFFu;
constexpr uint32_t kTagFar = 0xFFFEu;

struct LdsProbeArgs {
    const uint64_t *slots;
    uint32_t nbuckets;
    uint32_t slice;          // buckets per level-2 partition
    PartLayout L;            // level-2 layout (counted partitioning)
    uint32_t nparts;         // level-2 partitions
    // fill partitioning: partition p holds rows [p * cap, p * cap + min(fill[p], cap))
    const unsigned long long *fill;   // null: the counted layout L
    uint64_t cap;
    const int32_t *fcol[2];  // its hash and value columns
    uint32_t *next;          // partition counter (zeroed before the launch)
    TagGeom geom;            // tagged slices only
};

// Count-free partitioning: F regions of `cap` rows each, filled by
// reservations (one atomic per tile and partition) in any order, and an
// overflow area for the rows of a region that is full.  With cap set a few
// standard deviations above n / F, hash partitions of any input fit; a
// skewed input's excess goes to the overflow area, so every input is
// partitioned completely.
struct FillLayout {
    uint32_t F;
    uint32_t pad;
    uint64_t cap;                 // rows per region (multiple of 64)
    int32_t *col[2];              // hashes, values: F regions, 64 rows of padding, the overflow area, 64 more
    unsigned long long *fill;     // [F] rows reserved in each region (may pass cap)
    unsigned long long *ofill;    // rows reserved in the overflow area
    uint64_t ostart;              // first row of the overflow area: F * cap + 64
    uint64_t ocap;                // its rows
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
// (-1, -1)) for key (recovered from a row's hash with fmix32_inv): number of
// hitting slots and the sum of their payloads.  The empty pattern is
// excluded from the compare, so key -1 is handled.
__device__ __forceinline__ void bucket_scan(const v4u (&q)[4], uint32_t key, uint32_t &hits, int64_t &psum) {
    uint32_t h = 0;
    int64_t t = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t sk = (s & 1) ? q[s >> 1].z : q[s >> 1].x;
        const uint32_t sp = (s & 1) ? q[s >> 1].w : q[s >> 1].y;
        const bool hit = sk == key && !(sk == ~0u && sp == ~0u);
        h += hit ? 1u : 0u;
        t += hit ? (int64_t)(int32_t)sp : 0;
    }
    psum += t;
    hits += h;
}

__device__ __forceinline__ bool bucket_full(const v4u (&q)[4]) { return !(q[3].z == ~0u && q[3].w == ~0u); }

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// A key slice in LDS is stored keys first: per bucket its 8 keys (hashed,
// in two 16-byte quads), its 8 payloads, and its fill count c.  The build fills a
// bucket's slots strictly in order (a slot is claimed only after every lower
// slot was seen occupied, and slots are never freed), so the occupied slots
// are exactly [0, c): slot s hits key iff s < c and its key equals key (an
// empty slot never hits, whatever the key), and the bucket is full iff
// c == 8.  A probe row then reads 32 bytes of keys and 1 byte of count, and a
// payload only when it hits, instead of the 64-byte bucket.  NB is the slice
// size in buckets (the probe's 2048, the group-by probe's 1024, which leaves
// LDS for its group table).
template <uint32_t NB>
struct SliceT {
    static constexpr uint32_t kBuckets = NB;
    v4u key[(NB + kSliceExtra) * 2];
    uint32_t pay[(NB + kSliceExtra) * 8 + 1];   // the last word is 0: read for a row without a hit
    uint8_t cnt[NB + kSliceExtra];
    static constexpr uint32_t kZero = (NB + kSliceExtra) * 8;

    // one bucket's keys and fill count, as read for a probe
    struct Bucket {
        v4u k0, k1;
        uint32_t c;
    };
    __device__ __forceinline__ Bucket fetch(uint32_t idx) const;
    // slot mask of the hits of the row of hash h at chain offset j
    __device__ __forceinline__ static uint32_t hits(const Bucket &b, uint32_t h, uint32_t j, const TagGeom &g);
    __device__ __forceinline__ static bool full(const Bucket &b) { return b.c == 8; }
    // whether a row at chain offset j is probed in LDS (always)
    __device__ __forceinline__ static bool in_range(uint32_t, const TagGeom &) { return true; }
    // the bucket after b, and whether a chain from b may go on to it (not
    // when it is b itself: a one-bucket table)
    __device__ __forceinline__ static uint32_t next(uint32_t b, uint32_t nb) { return b + 1 == nb ? 0 : b + 1; }
    __device__ __forceinline__ static bool moves(uint32_t b, uint32_t nb) { return nb > 1; }
    __device__ __forceinline__ static uint32_t home(uint32_t h, uint32_t nb, const TagGeom &) {
        return bucket_of_hash(h, nb);
    }
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

template <uint32_t NB>
__device__ __forceinline__ typename SliceT<NB>::Bucket SliceT<NB>::fetch(uint32_t idx) const {
    return Bucket{key[key_at(idx, 0)], key[key_at(idx, 1)], cnt[idx]};
}

template <uint32_t NB>
__device__ __forceinline__ uint32_t SliceT<NB>::hits(const Bucket &b, uint32_t h, uint32_t, const TagGeom &) {
    return slice_hits(b.k0, b.k1, b.c, h);
}

// A tagged slice: per bucket its 8 slot tags in 16 bytes (dword w holds
// slot w in its low half and slot w + 4 in its high half) and its 8
// payloads.  The build fills a bucket's slots in order and never frees one,
// so the bucket is full exactly when slot 7 is not empty.
template <uint32_t NB>
struct SliceTag {
    static constexpr uint32_t kBuckets = NB;
    v4u tag[NB + kSliceExtra];
    uint32_t pay[(NB + kSliceExtra) * 8 + 1];   // the last word is 0: read for a row without a hit
    static constexpr uint32_t kZero = (NB + kSliceExtra) * 8;

    struct Bucket {
        v4u t;
    };
    __device__ __forceinline__ Bucket fetch(uint32_t idx) const { return Bucket{tag[idx]}; }
    // the row's tag against the 8 slot tags: xor, then per 16-bit half
    // 1 - x saturated at 0 (1 exactly when the tags are equal)
    __device__ __forceinline__ static uint32_t hits(const Bucket &b, uint32_t h, uint32_t j, const TagGeom &g) {
        const uint32_t r = (j << g.L) | (h & g.lowmask);
        const uint32_t R = r | (r << 16);
        const u16x2 one = {1, 1};
        const uint32_t d0 = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(one, __builtin_bit_cast(u16x2, b.t.x ^ R)));
        const uint32_t d1 = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(one, __builtin_bit_cast(u16x2, b.t.y ^ R)));
        const uint32_t d2 = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(one, __builtin_bit_cast(u16x2, b.t.z ^ R)));
        const uint32_t d3 = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(one, __builtin_bit_cast(u16x2, b.t.w ^ R)));
        const uint32_t m = ((d3 << 1 | d2) << 1 | d1) << 1 | d0;   // slot w at bit w, slot w + 4 at bit 16 + w
        return (m & 0xFu) | (m >> 12);                                // bits 12-15 of m are 0
    }
    __device__ __forceinline__ static bool full(const Bucket &b) { return b.t.w < (kTagEmpty << 16); }
    __device__ __forceinline__ static bool in_range(uint32_t j, const TagGeom &g) { return j <= g.jmax; }
    // nb is a power of two >= 2^18
    __device__ __forceinline__ static uint32_t next(uint32_t b, uint32_t nb) { return (b + 1) & (nb - 1); }
    __device__ __forceinline__ static bool moves(uint32_t, uint32_t) { return true; }
    // bucket_of_hash for nb = 2^(32 - L): the top bits
    __device__ __forceinline__ static uint32_t home(uint32_t h, uint32_t, const TagGeom &g) { return h >> g.L; }
};

__device__ __forceinline__ uint32_t lowest_slot(uint32_t hm) { return hm ? (uint32_t)__builtin_ctz(hm) : 0u; }

// What a probe does with its matches.  AccSum: COUNT and SUM(val + payload)
// in registers (the join aggregate, hashjoin.cu:135-182 with
// aggregation.cu:27-66); AccGroup: GROUP BY payload, COUNT and SUM(val),
// in a per-block LDS table of NG groups flushed with global atomics (groups
// >= ngroups are dropped, as mgv_hashjoin_probe_groupby drops them).
struct LdsAccArgs {
    mgv_result *res;       // AccSum
    uint64_t *gcnt;        // AccGroup
    int64_t *gsum;
    uint32_t ngroups;
};

struct AccSum {
    struct Smem {};
    uint64_t cnt = 0, sum = 0;   // sum: two's-complement int64
    __device__ __forceinline__ void open(Smem *, const LdsAccArgs &) {}
    // hit mask hm of LDS bucket idx (p1 = payload of its lowest hit, 0
    // without a hit)
    template <class S>
    __device__ __forceinline__ void lds(const S &sl, uint32_t idx, uint32_t hm, int32_t p1, int32_t val) {
        // branch-free for the common rows (at most one hit): count, and the
        // lowest hit's payload plus the row value
        const uint32_t h = (uint32_t)__builtin_popcount(hm);
        cnt += h;
        sum += (uint64_t)((int64_t)p1 + (int64_t)(hm ? val : 0));
        if (ballot(h > 1)) {
            int64_t t = 0;
            for (uint32_t m = hm & (hm - 1); m; m &= m - 1)
                t += (int64_t)(int32_t)sl.pay[idx * 8 + __builtin_ctz(m)] + (int64_t)val;
            sum += (uint64_t)t;
        }
    }
    // a global-memory bucket of a row whose hash is h_key
    __device__ __forceinline__ void bucket(const v4u (&q)[4], uint32_t h_key, int32_t val) {
        uint32_t hits = 0;
        int64_t psum = 0;
        bucket_scan(q, fmix32_inv(h_key), hits, psum);
        if (hits) {
            cnt += hits;
            sum += (uint64_t)(psum + (int64_t)hits * (int64_t)val);
        }
    }
    __device__ __forceinline__ void bucket_wave(bool live, const v4u (&q)[4], uint32_t h_key, int32_t val) {
        if (live) bucket(q, h_key, val);
    }
    // every wave of the block calls close
    __device__ __forceinline__ void close(const LdsAccArgs &a) {
        cnt = wave_sum_u64(cnt);
        sum = wave_sum_u64(sum);
        if (lane_id() == 0) {
            if (cnt) atomic_add_global((unsigned long long *)&a.res->count, (unsigned long long)cnt);
            if (sum) atomic_add_global((unsigned long long *)&a.res->sum, (unsigned long long)sum);
        }
    }
};

template <uint32_t NG>
struct AccGroup {
    struct Smem {
        uint32_t cnt[NG];
        unsigned long long sum[NG];
    };
    Smem *sm = nullptr;
    uint32_t ngroups = 0;
    __device__ __forceinline__ void add(uint32_t g, int32_t v) {
        if (g < ngroups) {
            atomic_add_lds(&sm->cnt[g], 1u);
            atomic_add_lds(&sm->sum[g], (unsigned long long)(int64_t)v);
        }
    }
    // the whole block calls open
    __device__ __forceinline__ void open(Smem *s, const LdsAccArgs &a) {
        sm = s;
        ngroups = a.ngroups < NG ? a.ngroups : NG;
        for (uint32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
            sm->cnt[g] = 0;
            sm->sum[g] = 0;
        }
        __syncthreads();
    }
    // One row per lane with `on`, called by every lane of the wave: when at
    // least kHotLanes lanes add to the same group (the lowest active lane's),
    // their values are summed across the wave and added with one atomic
    // pair -- rows partitioned by key arrive in runs of one hot key, whose
    // per-lane atomics on one LDS word would serialise the wave.
    static constexpr uint32_t kHotLanes = 16;
    __device__ __forceinline__ void add_wave(bool on, uint32_t g, int32_t v) {
        const bool ok = on && g < ngroups;
        const uint64_t act = ballot(ok);
        if (act == 0) return;
        const uint32_t first = (uint32_t)__builtin_ctzll(act);
        const uint32_t g0 = (uint32_t)__builtin_amdgcn_readlane((int)g, (int)first);
        const bool in0 = ok && g == g0;
        const uint64_t same = ballot(in0);
        bool mine = ok;
        if (popc64(same) >= kHotLanes) {
            const uint64_t t = wave_sum_u64(in0 ? (uint64_t)(int64_t)v : 0ull);
            if (lane_id() == first) {
                atomic_add_lds(&sm->cnt[g0], popc64(same));
                atomic_add_lds(&sm->sum[g0], (unsigned long long)t);
            }
            mine = ok && !in0;
        }
        if (mine) {
            atomic_add_lds(&sm->cnt[g], 1u);
            atomic_add_lds(&sm->sum[g], (unsigned long long)(int64_t)v);
        }
    }
    // called by every lane of the wave
    template <class S>
    __device__ __forceinline__ void lds(const S &sl, uint32_t idx, uint32_t hm, int32_t p1, int32_t val) {
        add_wave(hm != 0, (uint32_t)p1, val);
        if (ballot(__builtin_popcount(hm) > 1)) {
            for (uint32_t m = hm & (hm - 1); m; m &= m - 1) add(sl.pay[idx * 8 + __builtin_ctz(m)], val);
        }
    }
    // a global-memory bucket of a row whose hash is h_key (any lane set)
    __device__ __forceinline__ void bucket(const v4u (&q)[4], uint32_t h_key, int32_t val) {
        const uint32_t key = fmix32_inv(h_key);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t sk = (s & 1) ? q[s >> 1].z : q[s >> 1].x;
            const uint32_t sp = (s & 1) ? q[s >> 1].w : q[s >> 1].y;
            if (sk == key && !(sk == ~0u && sp == ~0u)) add(sp, val);
        }
    }
    // the same, called by every lane of the wave (rows of lanes with
    // `live`): the first hitting slot through add_wave, any further one (a
    // duplicate key) per lane
    __device__ __forceinline__ void bucket_wave(bool live, const v4u (&q)[4], uint32_t h_key, int32_t val) {
        const uint32_t key = fmix32_inv(h_key);
        uint32_t hm = 0, p0 = 0;   // hit mask, payload of the lowest hit
#pragma unroll
        for (int s = 7; s >= 0; --s) {
            const uint32_t sk = (s & 1) ? q[s >> 1].z : q[s >> 1].x;
            const uint32_t sp = (s & 1) ? q[s >> 1].w : q[s >> 1].y;
            const bool hit = live && sk == key && !(sk == ~0u && sp == ~0u);
            hm |= (uint32_t)hit << s;
            p0 = hit ? sp : p0;
        }
        add_wave(hm != 0, p0, val);
        if (ballot((hm & (hm - 1)) != 0)) {
            for (uint32_t m = hm & (hm - 1); m; m &= m - 1) {
                const uint32_t s = (uint32_t)__builtin_ctz(m);
                add((s & 1) ? q[s >> 1].w : q[s >> 1].y, val);
            }
        }
    }
    // the whole block calls close
    __device__ __forceinline__ void close(const LdsAccArgs &a) {
        __syncthreads();
        for (uint32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
            const uint32_t c = sm->cnt[g];
            if (c) {
                atomic_add_global((unsigned long long *)&a.gcnt[g], (unsigned long long)c);
                atomic_add_global((unsigned long long *)&a.gsum[g], sm->sum[g]);
            }
        }
    }
};

// Rows whose chain continues past their home bucket (about 5 % at load
// factor 0.5, but some row of nearly every wave) are not followed on the
// spot, which would cost the whole wave a scan per row: they are appended to
// a per-wave queue in LDS {key, value, bucket} and probed 64 at a time, one
// per lane, once the queue holds a full wave's worth (and at the end).
constexpr uint32_t kQueue = 128;   // entries per wave (a power of two)

// An entry's bkt is the bucket to probe, or (bit 31 set) the bucket whose
// successor is to be probed: a row whose LDS home bucket was full is queued
// without computing the successor (tables of the LDS path have at most
// 32768 slices of 2048 buckets, 2^26 buckets, so bit 31 is free).
constexpr uint32_t kAfter = 1u << 31;

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
        const uint32_t pos =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, tail)) &
            (kQueue - 1);
        Q.key[pos] = key;
        Q.val[pos] = val;
        Q.bkt[pos] = bkt;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)(tail + (uint32_t)popc64(bal)));
}

// Probes up to 64 queued entries (one per lane) and re-queues the chains
// that go on; returns the new head and, through tail, the new tail.  A
// bucket inside the loaded range is read from LDS (for a tagged slice: while
// the chain offset fits a tag), any other from the table.
template <class S, class Acc>
__device__ __forceinline__ uint32_t queue_round(WaveQueue &Q, uint32_t head, uint32_t &tail, const S &sl,
                                                const uint64_t *slots, uint32_t nb, uint32_t lo, uint32_t nload,
                                                const TagGeom &g, Acc &acc) {
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
        const uint32_t raw = Q.bkt[pos];
        bkt = (raw & kAfter) ? S::next(raw & ~kAfter, nb) : raw;
    }
    wave_sync();   // every lane has read its entry before slots are reused
    const uint32_t home = S::home(key, nb, g);
    const uint32_t idx = bkt - lo;
    const uint32_t j = bkt - home;   // the chain offset (huge for a chain that wrapped: never in LDS)
    const bool in_lds = mine && idx < nload && S::in_range(j, g);
    bool full = false;
    uint32_t hm = 0, ii = 0;
    if (in_lds) {
        const typename S::Bucket B = sl.fetch(idx);
        hm = S::hits(B, key, j, g);
        ii = idx;
        full = S::full(B);
    }
    acc.lds(sl, ii, hm, (int32_t)sl.pay[hm ? ii * 8 + (uint32_t)__builtin_ctz(hm) : S::kZero], val);
    if (mine && !in_lds) {
        const v4u *gb = reinterpret_cast<const v4u *>(slots + (uint64_t)bkt * 8);
        v4u q[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) q[w] = load_global(gb + w);
        acc.bucket(q, key, val);
        full = bucket_full(q);
    }
    // a chain ends at a bucket with an empty slot, or after visiting all nb
    // buckets once (when its next bucket is its home bucket again)
    const uint32_t nxt = (bkt + 1 == nb) ? 0 : bkt + 1;
    const bool more = mine && full && nxt != home;
    head += take;
    tail = queue_push(Q, tail, more, key, val, nxt);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)head);
}

// A lane's four rows: the four home buckets are read from LDS together and
// compared, then the four lowest-hit payloads are read together; rows whose
// bucket is full continue at the next bucket, and rows whose home bucket
// lies outside the slice start at their home bucket, through the wave's
// queue.
template <class S, class Acc>
__device__ __forceinline__ void lds_probe_quad(const v4i &k, const v4i &v, uint32_t m, const S &sl,
                                               const uint64_t *slots, uint32_t nb, uint32_t lo, uint32_t nload,
                                               const TagGeom &g, Acc &acc, WaveQueue &Q, uint32_t &head,
                                               uint32_t &tail) {
    uint32_t b[4], ii[4];
    bool ins[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        b[j] = S::home((uint32_t)q_get(k, j), nb, g);
        const uint32_t idx = b[j] - lo;
        ins[j] = ((m >> j) & 1u) && idx < nload;
        ii[j] = ins[j] ? idx : 0u;
    }
    typename S::Bucket B[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) B[j] = sl.fetch(ii[j]);
    // every lane compares (a lane without a row reads bucket 0), then masks
    uint32_t hm[4];
    int32_t p1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t x = S::hits(B[j], (uint32_t)q_get(k, j), 0u, g);
        hm[j] = ins[j] ? x : 0u;
    }
    // the lowest hit's payload (the slice's zero word when there is none)
#pragma unroll
    for (int j = 0; j < 4; ++j)
        p1[j] = (int32_t)sl.pay[hm[j] ? ii[j] * 8 + (uint32_t)__builtin_ctz(hm[j]) : S::kZero];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        acc.lds(sl, ii[j], hm[j], p1[j], q_get(v, j));
        const bool live = (m >> j) & 1u;
        // a row in the slice goes on after its bucket when the bucket is
        // full, any other live row starts at its home bucket (bitwise:
        // lane-mask operations)
        const uint32_t from = ins[j] ? (b[j] | kAfter) : b[j];
        const bool go = (ins[j] & S::full(B[j]) & S::moves(b[j], nb)) | (!ins[j] & live);
        tail = queue_push(Q, tail, go, (uint32_t)q_get(k, j), q_get(v, j), from);
        // fewer than kWave entries were queued before this push and at most
        // kWave were added, so the queue never holds more than kQueue - 1
        while (tail - head >= kWave) head = queue_round(Q, head, tail, sl, slots, nb, lo, nload, g, acc);
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

// Copies slice buckets [lo, lo + nload) of the table into the slice, one
// bucket per thread: keys, payloads and the fill count (the first empty
// slot, or 8).
template <uint32_t NB>
__device__ __forceinline__ void load_slice(SliceT<NB> &sl, const uint64_t *slots, uint32_t lo, uint32_t nload,
                                           uint32_t, const TagGeom &) {
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
        sl.key[key_at(i, 0)] = v4u{fmix32(q0.x), fmix32(q0.z), fmix32(q1.x), fmix32(q1.z)};
        sl.key[key_at(i, 1)] = v4u{fmix32(q2.x), fmix32(q2.z), fmix32(q3.x), fmix32(q3.z)};
        v4u *pay4 = reinterpret_cast<v4u *>(sl.pay);
        pay4[i * 2] = v4u{q0.y, q0.w, q1.y, q1.w};
        pay4[i * 2 + 1] = v4u{q2.y, q2.w, q3.y, q3.w};
        sl.cnt[i] = (uint8_t)c;
    }
}

// The tagged copy: slot s < c of global bucket b with key k gets the tag of
// fmix32(k) at chain offset j = (b - home) mod nb (kTagFar when j does not
// fit), slots from c on kTagEmpty.
template <uint32_t NB>
__device__ __forceinline__ void load_slice(SliceTag<NB> &sl, const uint64_t *slots, uint32_t lo, uint32_t nload,
                                           uint32_t nb, const TagGeom &tg) {
    for (uint32_t i = threadIdx.x; i < nload; i += kLdsBlock) {
        const uint32_t b = lo + i;
        const v4u *g = reinterpret_cast<const v4u *>(slots + (uint64_t)b * 8);
        const v4u q0 = load_global(g), q1 = load_global(g + 1), q2 = load_global(g + 2), q3 = load_global(g + 3);
        const v4u qs[4] = {q0, q1, q2, q3};
        uint32_t c = 8;
#pragma unroll
        for (int s = 7; s >= 0; --s) {
            const v4u &q = qs[s >> 1];
            const bool empty = (s & 1) ? (q.z == ~0u && q.w == ~0u) : (q.x == ~0u && q.y == ~0u);
            c = empty ? (uint32_t)s : c;
        }
        uint32_t t[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const v4u &q = qs[s >> 1];
            const uint32_t hk = fmix32((s & 1) ? q.z : q.x);
            const uint32_t j = (b - (hk >> tg.L)) & (nb - 1u);
            const uint32_t tag = j <= tg.jmax ? (j << tg.L) | (hk & tg.lowmask) : kTagFar;
            t[s] = (uint32_t)s < c ? tag : kTagEmpty;
        }
        sl.tag[i] = v4u{t[0] | t[4] << 16, t[1] | t[5] << 16, t[2] | t[6] << 16, t[3] | t[7] << 16};
        v4u *pay4 = reinterpret_cast<v4u *>(sl.pay);
        pay4[i * 2] = v4u{q0.y, q0.w, q1.y, q1.w};
        pay4[i * 2 + 1] = v4u{q2.y, q2.w, q3.y, q3.w};
    }
}

template <class Slice, class Acc>
__global__ __launch_bounds__(kLdsBlock) void k_lds_probe(LdsProbeArgs a, LdsAccArgs aa) {
    __shared__ __align__(16) Slice S;
    __shared__ WaveQueue s_queue[kLdsBlock / kWave];
    __shared__ __align__(16) typename Acc::Smem s_acc;
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    WaveQueue &Q = s_queue[t / kWave];
    const uint32_t nb = a.nbuckets;
    const int32_t *K = a.fill ? a.fcol[0] : a.L.col[0];
    const int32_t *V = a.fill ? a.fcol[1] : a.L.col[1];
    Acc acc;
    acc.open(&s_acc, aa);
    if (t == 0) S.pay[Slice::kZero] = 0u;   // visible after the first barrier below
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t p = s_part;
        __syncthreads();   // s_part is rewritten next round; the slice is free again
        if (p >= a.nparts) break;
        uint64_t r0, r1;
        if (a.fill) {
            r0 = (uint64_t)p * a.cap;
            const unsigned long long f = a.fill[p];
            r1 = r0 + (f < a.cap ? f : a.cap);
        } else {
            r0 = a.L.pstart[p];
            r1 = r0 + a.L.ptotal[p];
        }
        if (r0 == r1) continue;
        const uint32_t lo = p * a.slice;
        const uint32_t nload = min(a.slice + kSliceExtra, nb - lo);   // a.slice <= NB
        // the first rows' loads overlap the slice copy
        v4i k[kAhead], v[kAhead];
        uint32_t m[kAhead];
        uint64_t i0 = r0 + 4ull * t;
#pragma unroll
        for (int x = 0; x < kAhead; ++x) {
            const uint64_t ix = i0 + (uint64_t)x * kQuadStride;
            m[x] = load_rows4(K, V, ix < r1 ? ix : r0, ix < r1 ? r1 : r0, k[x], v[x]);
        }
        load_slice(S, a.slots, lo, nload, nb, a.geom);
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
                lds_probe_quad(k[x], v[x], m[x], S, a.slots, nb, lo, nload, a.geom, acc, Q, head, tail);
#pragma unroll
            for (int x = 0; x < kAhead; ++x) {
                k[x] = kn[x];
                v[x] = vn[x];
                m[x] = mn[x];
            }
            i0 = i1;
        }
        // the rest of the queue, before the slice is replaced
        while (tail != head) head = queue_round(Q, head, tail, S, a.slots, nb, lo, nload, a.geom, acc);
        __syncthreads();
    }
    acc.close(aa);
}

// ------------------------------------------------------------------ fill partitioning

constexpr int kFillBlock = 1024;
constexpr int kFillTile = 8192;        // rows per tile of a scatter into up to 2048 partitions
constexpr int kFillTileWide = 16384;   // ... into up to 512 (the staging takes 128 KiB of LDS)
constexpr int kFillBlocksPerCU = 2;    // grid of the scatters: blocks per CU

// One tile of rows of a fill scatter's input: the rows of its segment from
// r0 (below r1), whose partitions are [base, base + S).
struct FillTile {
    const int32_t *k, *v, *pred;
    uint64_t r0, r1;
    uint32_t base;
};

// Scatters rows into a FillLayout by the top bits of their hash, tile by
// tile: each tile is counting-sorted in LDS by partition, every partition
// with rows in the tile reserves its run with one atomic on its fill count
// (the part of a run past the region's capacity goes to the overflow area),
// and the staged rows are written run by run.  Input: FROM_UP = false, n
// rows of key (hashed here with fmix32), val and an optional filter pred <=
// bound, and a tile spans all F partitions; FROM_UP = true, the regions of
// an upstream FillLayout U (rows already hashed; its overflow area is probed
// as it is and not passed on), whose partitions are coarser: region a holds
// the rows whose top log2(U.F) hash bits are a, so its tiles span only the
// S = F / U.F partitions [a S, a S + S), the per-tile arrays hold S entries
// (MAXS bounds S) and the runs are long.  The next tile's columns load while
// the current one is written; the write step issues a fixed number of stores
// per thread (lanes past the tile's last staged row repeat its store; an
// empty tile stores into the padding after the regions), so the compiler
// can wait for the loads without the stores.  The next tile's loads are
// issued after the run reservations (whose returned offsets are waited for
// in issue order with every earlier memory operation) and the staging.
template <bool FROM_UP, uint32_t MAXS, int TILE>
__global__ __launch_bounds__(kFillBlock) void k_fill_scatter(FillLayout O, FillLayout U, const int32_t *key,
                                                             const int32_t *val, const int32_t *pred, int32_t bound,
                                                             uint64_t n) {
    constexpr int BLOCK = kFillBlock;
    constexpr int Q = TILE / (4 * BLOCK);
    constexpr int R = TILE / BLOCK;
    constexpr int PER = MAXS > (uint32_t)BLOCK ? (int)(MAXS / BLOCK) : 1;
    static_assert(Q >= 1 && TILE % (4 * BLOCK) == 0, "fill scatter shape");
    __shared__ unsigned long long dbase[MAXS];   // slot i of local partition q -> region row dbase[q] + i
    __shared__ unsigned long long obase[MAXS];   // slot i past rlim[q] -> overflow row obase[q] + i
    __shared__ uint32_t rlim[MAXS];              // slots below rlim[q] fit the region
    __shared__ uint32_t tcnt[MAXS];
    __shared__ uint16_t toff[MAXS];
    __shared__ uint32_t wtot[BLOCK / kWave];
    __shared__ uint32_t s_total;
    __shared__ uint32_t s_tiles[FROM_UP ? BLOCK + 1 : 1];   // cumulative tiles per upstream region
    __shared__ uint32_t s_len[FROM_UP ? BLOCK : 1];         // rows of each upstream region
    __shared__ int32_t sk[TILE];
    __shared__ int32_t sv[TILE];
    const uint32_t F = O.F;
    const uint32_t S = FROM_UP ? F / U.F : F;   // partitions a tile spans
    const uint32_t t = threadIdx.x;
    const uint32_t per = (S + BLOCK - 1) / BLOCK;
    for (uint32_t q = t; q < S; q += BLOCK) tcnt[q] = 0;
    // the tiles of the input
    uint32_t ntiles;
    if (FROM_UP) {
        const uint32_t nseg = U.F;   // <= BLOCK
        uint64_t len = 0;
        if (t < nseg) {
            const unsigned long long f = U.fill[t];
            len = f < U.cap ? f : U.cap;
            s_len[t] = (uint32_t)len;
        }
        const uint32_t nt = t < nseg ? (uint32_t)((len + TILE - 1) / TILE) : 0u;
        const uint32_t before = scan_block_excl<BLOCK>(nt, wtot);
        if (t < nseg) s_tiles[t + 1] = before + nt;
        if (t == 0) s_tiles[0] = 0;
        __syncthreads();
        ntiles = s_tiles[nseg];
    } else {
        ntiles = (uint32_t)((n + TILE - 1) / TILE);
    }
    // the region of the block's current tile: tiles are taken in increasing
    // order, so it only moves forward (usually by at most one region)
    uint32_t seg = 0;
    auto tile_of = [&](uint32_t ti) {
        FillTile ft;
        if (FROM_UP) {
            // the region a with s_tiles[a] <= ti < s_tiles[a + 1]
            uint32_t a = seg;
            while (a + 1 < U.F && s_tiles[a + 1] <= ti) ++a;
            a = (uint32_t)__builtin_amdgcn_readfirstlane((int)a);
            seg = a;
            ft.k = U.col[0] + (uint64_t)a * U.cap;
            ft.v = U.col[1] + (uint64_t)a * U.cap;
            ft.r1 = s_len[a];
            ft.pred = nullptr;
            ft.r0 = (uint64_t)(ti - s_tiles[a]) * TILE;
            ft.base = a * S;
        } else {
            ft.k = key;
            ft.v = val;
            ft.pred = pred;
            ft.r0 = (uint64_t)ti * TILE;
            ft.r1 = n;
            ft.base = 0;
        }
        if (ft.r1 > ft.r0 + TILE) ft.r1 = ft.r0 + TILE;
        return ft;
    };
    v4i nk[Q], nv[Q], nc[Q];
    uint32_t nb[Q];
    auto load_tile = [&](const FillTile &ft) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint64_t i0 = ft.r0 + 4ull * t + (uint64_t)q * 4 * BLOCK;
            uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) m |= (uint32_t)(i0 + j < ft.r1) << j;
            nb[q] = m;
            nk[q] = load_quad(ft.k, i0, ft.r1);
            nv[q] = ft.v ? load_quad(ft.v, i0, ft.r1) : v4i{0, 0, 0, 0};
            nc[q] = ft.pred ? load_quad(ft.pred, i0, ft.r1) : v4i{0, 0, 0, 0};
        }
    };
    // the block's tiles: every gridDim.x-th tile, so the blocks running at
    // the same time work on neighbouring tiles (measured: 7 % faster at
    // level 2 than contiguous runs of tiles per block)
    uint32_t ti = blockIdx.x;
    const uint32_t step = gridDim.x;
    bool has_pred = false;
    uint32_t nbase = 0;   // first partition of the loaded tile
    if (ti < ntiles) {
        const FillTile ft = tile_of(ti);
        has_pred = ft.pred != nullptr;
        nbase = ft.base;
        load_tile(ft);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        asm volatile("" ::"v"(nk[q].x), "v"(nk[q].y), "v"(nk[q].z), "v"(nk[q].w), "v"(nv[q].x), "v"(nv[q].y),
                     "v"(nv[q].z), "v"(nv[q].w), "v"(nc[q].x), "v"(nc[q].y), "v"(nc[q].z), "v"(nc[q].w));
    __syncthreads();
    const uint64_t dummy = (uint64_t)F * O.cap + 32;   // padding row: stores of an empty tile
    for (; ti < ntiles; ti += step) {
        // this tile's rows as hashes, with their masks
        const uint32_t base = nbase;
        v4i ch[Q], cv[Q];
        uint32_t cm[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            uint32_t m = nb[q];
            if (has_pred) {
#pragma unroll
                for (int j = 0; j < 4; ++j) m &= ~((uint32_t)(q_get(nc[q], j) > bound) << j);
            }
            cm[q] = m;
            cv[q] = nv[q];
            if (FROM_UP) {
                ch[q] = nk[q];
            } else {
                ch[q] = v4i{(int32_t)fmix32((uint32_t)nk[q].x), (int32_t)fmix32((uint32_t)nk[q].y),
                            (int32_t)fmix32((uint32_t)nk[q].z), (int32_t)fmix32((uint32_t)nk[q].w)};
            }
        }
        // 1. rank every row within its partition
        uint32_t rank[Q][4];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (cm[q] >> j) & 1u;
                rank[q][j] = ok ? atomicAdd(&tcnt[part_of_hash(q_get(ch[q], j), F) - base], 1u) : 0u;
            }
        __syncthreads();
        // 2. owners: staging offsets, run reservations, overflow
        uint32_t c[PER];
        uint32_t mine = 0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t q = t * per + e;
            c[e] = (e < (int)per && q < S) ? tcnt[q] : 0u;
            mine += c[e];
        }
        uint32_t start = scan_block_excl<BLOCK>(mine, wtot);
        if (t == BLOCK - 1) s_total = start + mine;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t q = t * per + e;
            if (e < (int)per && q < S) {
                toff[q] = (uint16_t)start;
                if (c[e]) {
                    const uint32_t p = base + q;
                    const unsigned long long off = atomic_add_global(&O.fill[p], (unsigned long long)c[e]);
                    const uint64_t room = off < O.cap ? O.cap - off : 0;
                    const uint32_t fit = room < c[e] ? (uint32_t)room : c[e];
                    dbase[q] = (uint64_t)p * O.cap + off - start;
                    rlim[q] = start + fit;
                    if (fit < c[e]) {
                        const unsigned long long o =
                            atomic_add_global(O.ofill, (unsigned long long)(c[e] - fit));
                        obase[q] = O.ostart + o - (start + fit);
                    }
                }
                tcnt[q] = 0;
                start += c[e];
            }
        }
        __syncthreads();
        // 3. stage rows in partition order
        uint32_t pos[Q][4];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (cm[q] >> j) & 1u;
                pos[q][j] = toff[ok ? part_of_hash(q_get(ch[q], j), F) - base : 0u] + rank[q][j];
            }
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((cm[q] >> j) & 1u) {
                    sk[pos[q][j]] = q_get(ch[q], j);
                    sv[pos[q][j]] = q_get(cv[q], j);
                }
        // the next tile's loads: issued after the reservations (whose returns
        // would otherwise wait for them) and after this tile's rows are
        // staged (their registers are free again), in flight while the tile
        // is written
        if (ti + step < ntiles) {
            const FillTile ft = tile_of(ti + step);
            nbase = ft.base;
            load_tile(ft);
        }
        __syncthreads();
        // 4. write the runs: exactly R stores per column and thread
        const uint32_t total = s_total;
        constexpr int G = R < 8 ? R : (R > 8 ? 4 : 8);   // stores in flight per group: fewer for the wide tile
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += G) {
            int32_t k[G], v[G];
            uint32_t ii[G];
#pragma unroll
            for (int r = 0; r < G; ++r) {
                ii[r] = total ? min(t + (uint32_t)(r0 + r) * BLOCK, total - 1) : 0u;
                k[r] = sk[ii[r]];
                v[r] = sv[ii[r]];
            }
            uint64_t dst[G];
#pragma unroll
            for (int r = 0; r < G; ++r) {
                const uint32_t q = total ? part_of_hash(k[r], F) - base : 0u;
                dst[r] = !total ? dummy : (ii[r] < rlim[q] ? dbase[q] : obase[q]) + ii[r];
            }
#pragma unroll
            for (int r = 0; r < G; ++r) {
                O.col[0][dst[r]] = k[r];
                O.col[1][dst[r]] = v[r];
            }
        }
        __syncthreads();
    }
}

// Rows of a fill partitioning's overflow area (rows whose region was full):
// `*nrows` rows (at most ncap) of hashes H and values V.
struct HashedRows {
    const int32_t *H, *V;
    const unsigned long long *nrows;
    uint64_t ncap;
};

// Probes hashed rows against the global table: the overflow areas of both
// levels of a fill partitioning, one row per lane, its chain followed bucket
// by bucket as in mgv_hashjoin_probe.  The chain loop is wave-uniform (a
// lane whose chain ended stays in it without a row), so the accumulator can
// combine the lanes' rows of one group.
template <class Acc>
__global__ __launch_bounds__(256) void k_probe_hashed(const uint64_t *slots, uint32_t nb, HashedRows A, HashedRows B,
                                                      LdsAccArgs aa) {
    __shared__ __align__(16) typename Acc::Smem s_acc;
    Acc acc;
    acc.open(&s_acc, aa);
    const unsigned long long fa = A.nrows ? *A.nrows : 0ull, fb = B.nrows ? *B.nrows : 0ull;
    const uint64_t na = fa < A.ncap ? fa : A.ncap;
    const uint64_t n = na + (fb < B.ncap ? fb : B.ncap);
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < n; i0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = i0 + threadIdx.x;
        bool live = i < n;
        const uint32_t h = live ? (uint32_t)(i < na ? A.H[i] : B.H[i - na]) : 0u;
        const int32_t v = live ? (i < na ? A.V[i] : B.V[i - na]) : 0;
        const uint32_t home = bucket_of_hash(h, nb);
        uint32_t b = home;
        while (ballot(live)) {
            v4u q[4] = {};
            if (live) {
                const v4u *g = reinterpret_cast<const v4u *>(slots + (uint64_t)b * 8);
#pragma unroll
                for (int w = 0; w < 4; ++w) q[w] = load_global(g + w);
            }
            acc.bucket_wave(live, q, h, v);
            const bool full = bucket_full(q);
            b = (b + 1 == nb) ? 0 : b + 1;
            live = live && full && b != home;
        }
    }
    acc.close(aa);
}

// Rows per region of a fill partitioning of up to n rows into F regions:
// n / F plus eight standard deviations of a uniform hash partition's size,
// plus 256, in multiples of 64 rows.
static uint64_t fill_cap(uint64_t n, uint32_t F) {
    const double lam = (double)n / F;
    const uint64_t c = (uint64_t)(lam + 8.0 * __builtin_sqrt(lam)) + 256;
    return (c + 63) / 64 * 64;
}

// Lays out a fill partitioning of up to n rows (regions, padding, overflow
// area, padding; two columns) at ws + off; the counters fill (F words) and
// ofill are placed by the caller.  Returns the end offset.
static uint64_t fill_plan(uint64_t n, uint32_t F, char *ws, uint64_t off, unsigned long long *fill,
                          unsigned long long *ofill, FillLayout *L) {
    const uint64_t cap = fill_cap(n, F);
    const uint64_t rows = (uint64_t)F * cap + 64 + n + 64;
    const uint64_t cbytes = (rows * 4 + 255) / 256 * 256;
    if (L) {
        L->F = F;
        L->pad = 0;
        L->cap = cap;
        L->col[0] = (int32_t *)(ws + off);
        L->col[1] = (int32_t *)(ws + off + cbytes);
        L->fill = fill;
        L->ofill = ofill;
        L->ostart = (uint64_t)F * cap + 64;
        L->ocap = n;
    }
    return off + 2 * cbytes;
}

// ------------------------------------------------------------------ build

struct LdsBuildArgs {
    uint64_t *slots;
    uint32_t nbuckets;
    uint32_t slice;          // buckets per level-2 partition
    PartLayout L;            // level-2 layout (key, payload) of the counted partitioning
    uint32_t nparts;         // level-2 partitions
    // fill partitioning: partition p holds rows [p * cap, p * cap + min(fill[p], cap))
    // of fmix32(key) and payload; null: the counted layout L (keys as they are)
    const unsigned long long *fill;
    uint64_t cap;
    const int32_t *fcol[2];
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
template <uint32_t NB>
__global__ __launch_bounds__(kLdsBlock) void k_lds_build(LdsBuildArgs a) {
    __shared__ __align__(16) unsigned long long tab[NB * 8];
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    const bool hashed = a.fill != nullptr;
    const int32_t *K = hashed ? a.fcol[0] : a.L.col[0];
    const int32_t *V = hashed ? a.fcol[1] : a.L.col[1];
    uint64_t errs = 0;
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t p = s_part;
        __syncthreads();
        if (p >= a.nparts) break;
        uint64_t r0, r1;
        if (hashed) {
            r0 = (uint64_t)p * a.cap;
            const unsigned long long f = a.fill[p];
            r1 = r0 + (f < a.cap ? f : a.cap);
        } else {
            r0 = a.L.pstart[p];
            r1 = r0 + a.L.ptotal[p];
        }
        if (r0 == r1) continue;   // an untouched slice is not copied
        const uint32_t lo = p * a.slice;
        v4u *g = reinterpret_cast<v4u *>(a.slots + (uint64_t)lo * 8);
        v4u *l = reinterpret_cast<v4u *>(tab);
        for (uint32_t i = t; i < a.slice * 4; i += kLdsBlock) l[i] = load_global(g + i);
        __syncthreads();
        // r0 is a multiple of 64 rows; columns are padded past the partition
        // end (counted) or the region runs on to its capacity (fill)
        for (uint64_t i0 = r0 + 4ull * t; i0 < r1; i0 += 4ull * kLdsBlock) {
            const v4i k = load_stream4(K + i0);
            const v4i v = load_stream4(V + i0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (i0 + j >= r1) break;
                const uint32_t kh = (uint32_t)q_get(k, j), pay = (uint32_t)q_get(v, j);
                const uint32_t key = hashed ? fmix32_inv(kh) : kh;
                const uint64_t val = ((uint64_t)pay << 32) | key;
                if (val == kEmptySlot) {
                    ++errs;
                    continue;
                }
                // home buckets lie in the slice (bucket / slice = level-2
                // partition); anything else is finished in the global table
                const uint32_t idx =
                    (hashed ? bucket_of_hash(kh, a.nbuckets) : bucket_of((int32_t)key, a.nbuckets)) - lo;
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
// The rows of a fill partitioning's overflow areas (A, B: hashed keys) come
// after the list.
__global__ __launch_bounds__(256) void k_lds_build_overflow(LdsBuildArgs a, HashedRows A, HashedRows B) {
    OpBuild<256 / kWave> op;
    op.slots = a.slots;
    op.nbuckets = a.nbuckets;
    op.res = a.res;
    op.consume_open();
    const uint64_t nl = *a.ovf_count;
    const unsigned long long fa = A.nrows ? *A.nrows : 0ull, fb = B.nrows ? *B.nrows : 0ull;
    const uint64_t na = fa < A.ncap ? fa : A.ncap, nbr = fb < B.ncap ? fb : B.ncap;
    const uint64_t n = nl + na + nbr;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        if (i < nl) {
            op.insert((uint32_t)a.ovf_key[i], (uint32_t)a.ovf_pay[i]);
        } else {
            const uint64_t j = i - nl;
            const HashedRows &h = j < na ? A : B;
            const uint64_t r = j < na ? j : j - na;
            op.insert(fmix32_inv((uint32_t)h.H[r]), (uint32_t)h.V[r]);
        }
    }
    op.consume_close();
}

// Slices of the count-free (fill) partitionings: level 2 cuts each of the F1
// level-1 regions into at most kMaxFanoutWide / 2 slices (the largest
// per-tile array of k_fill_scatter), and at most kMaxSlicesFill in all.  The
// counted partitioning stops at kMaxFanoutWide slices.
constexpr uint32_t kMaxSlicesFill = 32768;
constexpr uint32_t kMaxSlicesPerRegion = kMaxFanoutWide / 2;

// Level-2 fan-out for a table of nb buckets: the smallest power of two
// F2 >= 2 * F1, at most maxf, that divides nb into slices of at most smax
// buckets.  When F1 alone already cuts the table into such slices (or
// F1 > kMaxFanout asks for one level), there is one level only: F2 = F1.
// Returns false (and sets nothing) when there is no such F2.
static bool fanout_of(uint32_t nb, uint32_t F1, uint32_t smax, uint32_t maxf, uint32_t *F2) {
    if (F1 > kMaxFanout || (nb % F1 == 0 && nb / F1 <= smax && nb / F1 > 0)) {
        if (nb % F1 != 0 || nb / F1 > smax || nb / F1 == 0) return false;
        *F2 = F1;
        return true;
    }
    uint64_t f = 2ull * F1;
    while (f <= maxf && (nb % f != 0 || nb / f > smax)) f <<= 1;
    if (f > maxf || nb % f != 0 || nb / f > smax || nb / f == 0) return false;
    *F2 = (uint32_t)f;
    return true;
}

static int lds_fanout(uint32_t nb, uint32_t F1, uint32_t *F2, uint32_t smax = kSliceMax,
                      uint32_t maxf = kMaxFanoutWide) {
    if (fanout_of(nb, F1, smax, maxf, F2)) return MGV_OK;
    if (F1 > kMaxFanout)
        return fail(MGV_E_INVALID, "LDS probe: one level of %u partitions does not cut %u buckets into slices "
                    "of <= %u buckets", F1, nb, smax);
    return fail(MGV_E_INVALID, "LDS probe: %u buckets do not split into <= %u slices of <= %u buckets", nb, maxf,
                smax);
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

// Workspace of the fill-partitioned probe: 256 bytes of counters (the
// partition counter of the probe, the two overflow counts), the fill counts
// of both levels, then the level-1 layout (two levels only) and the level-2
// layout.
static int fill_probe_plan(uint64_t n, uint32_t nb, uint32_t F1, void *ws, uint64_t ws_bytes, FillLayout *L1,
                           FillLayout *L2, uint32_t *F2, uint64_t *need, uint32_t smax = kSliceMax) {
    if (F1 < 2 || F1 > kMaxFanoutWide || (F1 & (F1 - 1)))
        return fail(MGV_E_INVALID, "LDS probe: level-1 fan-out %u not a power of two in [2,%u]", F1, kMaxFanoutWide);
    int rc = lds_fanout(nb, F1, F2, smax, kMaxSlicesFill);
    if (rc) return rc;
    const bool one = *F2 == F1;
    if (!one && *F2 / F1 > kMaxSlicesPerRegion)
        return fail(MGV_E_INVALID, "LDS probe: %u buckets need %u slices per level-1 region (at most %u); raise the "
                    "fan-out", nb, *F2 / F1, kMaxSlicesPerRegion);
    char *b = (char *)ws;
    unsigned long long *ctr = (unsigned long long *)ws;
    const uint64_t counters = (256 + (uint64_t)(F1 + *F2) * 8 + 255) / 256 * 256;
    uint64_t off = counters;
    if (!one) off = fill_plan(n, F1, b, off, ws ? ctr + 32 : nullptr, ws ? ctr + 1 : nullptr, ws ? L1 : nullptr);
    off = fill_plan(n, *F2, b, off, ws ? ctr + 32 + F1 : nullptr, ws ? ctr + 2 : nullptr, ws ? L2 : nullptr);
    if (need) *need = off;
    if (!ws) return MGV_OK;
    if (ws_bytes < off) return fail(MGV_E_CAPACITY, "LDS probe: workspace too small");
    if (one) L1->F = 0;
    return MGV_OK;
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

// Slices of the fill-partitioned build: 1024 buckets (64 KiB of slots, so
// two blocks share a CU and one inserts while the other copies), or 2048
// when 1024-bucket slices would be too many.
constexpr uint32_t kBuildSlice = 1024;
static uint32_t build_slice(uint32_t nb, uint32_t F1) {
    uint32_t F2 = 0;
    if (F1 >= 2 && !(F1 & (F1 - 1)) && fanout_of(nb, F1, kBuildSlice, kMaxSlicesFill, &F2) &&
        (F2 == F1 || F2 / F1 <= kMaxSlicesPerRegion))
        return kBuildSlice;
    return kSliceMax;
}

// Fill-partitioned build workspace: the fill probe's (counters, both
// partitionings of n rows, slices of build_slice buckets) followed by the
// chain-overflow list (n keys, n payloads); the list's count is counter
// word 3.
static int build_fill_plan(uint64_t n, uint32_t nb, uint32_t F1, void *ws, uint64_t ws_bytes, FillLayout *L1,
                           FillLayout *L2, uint32_t *F2, int32_t **okey, int32_t **opay, uint64_t *need) {
    uint64_t base = 0;
    const uint32_t smax = build_slice(nb, F1);
    int rc = fill_probe_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, F2, &base, smax);
    if (rc) return rc;
    base = (base + 255) / 256 * 256;
    const uint64_t ob = (n * 4 + 255) / 256 * 256;
    if (need) *need = base + 2 * ob;
    if (!L1) return MGV_OK;
    if (!ws || ws_bytes < base + 2 * ob) return fail(MGV_E_CAPACITY, "LDS build: workspace too small");
    if ((rc = fill_probe_plan(n, nb, F1, ws, base, L1, L2, F2, nullptr, smax))) return rc;
    *okey = (int32_t *)((char *)ws + base);
    *opay = (int32_t *)((char *)ws + base + ob);
    return MGV_OK;
}

// Workspace of either build (the counted one only where its fan-out limit
// allows it).
static int build_ws_need(uint64_t n, uint32_t nb, uint32_t F1, uint64_t *need) {
    uint32_t F2 = 0;
    uint64_t a = 0, b = 0;
    int rc = build_fill_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, &F2, nullptr, nullptr, &b);
    if (rc) return rc;
    if (fanout_of(nb, F1, kSliceMax, kMaxFanoutWide, &F2) &&
        (rc = build_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, &F2, nullptr, nullptr, &a)))
        return rc;
    *need = a > b ? a : b;
    return MGV_OK;
}

// Workspace of either partitioning (a caller's workspace serves both; a
// table too large for the counted one gets the fill partitioning's).
static int probe_ws_need(uint64_t n, uint32_t nb, uint32_t F1, uint64_t *need) {
    uint32_t F2 = 0;
    uint64_t a = 0, b = 0;
    int rc = fill_probe_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, &F2, &b, kProbeSlice);
    if (rc) return rc;
    if (F1 >= 2 && F1 <= kMaxFanoutWide && !(F1 & (F1 - 1)) && fanout_of(nb, F1, kSliceMax, kMaxFanoutWide, &F2) &&
        (rc = lds_plan(n, nb, F1, nullptr, 0, nullptr, nullptr, &F2, &a)))
        return rc;
    *need = a > b ? a : b;
    return MGV_OK;
}

// Tag geometry of a table of nbuckets = 2^B buckets, 18 <= B <= 31 (false
// for any other table, which keeps key slices).
static bool tag_geometry(uint32_t nbuckets, TagGeom *g) {
    if (nbuckets & (nbuckets - 1)) return false;
    const uint32_t B = (uint32_t)__builtin_ctz(nbuckets);
    if (B < 18 || B > 31) return false;
    g->L = 32 - B;
    g->lowmask = (1u << g->L) - 1u;
    g->jmax = (1u << (B - 16)) - 2u;
    return true;
}

// CUs of the current device (256 if unknown).
static uint32_t device_cus() {
    DeviceState *ds = current_state();
    return (ds && ds->cus > 0) ? (uint32_t)ds->cus : 256u;
}

// Launches the count-free partitioning of n rows (key hashed with fmix32,
// val, optional filter pred <= bound) into the layouts of fill_probe_plan:
// level 1 into L1's F1 regions, then each region into F2 / F1 regions of L2;
// or, when L1.F == 0, one level straight into L2.
static int fill_partition(const FillLayout &L1, const FillLayout &L2, uint32_t F1, uint32_t F2, const int32_t *keys,
                          const int32_t *vals, const int32_t *pred, int32_t bound, uint64_t n, hipStream_t s) {
    int rc;
    const uint64_t gmax = (uint64_t)kFillBlocksPerCU * device_cus();
    // blocks for a scatter of up to `rows` rows in tiles of `tile` (plus one
    // partial tile per upstream region)
    auto grid_for = [&](uint64_t rows, int tile, uint32_t regions) {
        const uint64_t t = (rows + tile - 1) / tile + regions;
        return (uint32_t)(t < gmax ? t : gmax);
    };
    if (L1.F) {
        hipLaunchKernelGGL((k_fill_scatter<false, kMaxFanout, kFillTileWide>), dim3(grid_for(n, kFillTileWide, 0)),
                           dim3(kFillBlock), 0, s, L1, L1, keys, vals, pred, bound, n);
        if ((rc = check_launch("k_fill_scatter (level 1)"))) return rc;
        // a level-2 tile spans F2 / F1 partitions: small per-tile arrays leave
        // LDS for the wide tile
        const uint32_t S = F2 / F1;
        if (S <= 64)
            hipLaunchKernelGGL((k_fill_scatter<true, 64, kFillTileWide>), dim3(grid_for(n, kFillTileWide, F1)),
                               dim3(kFillBlock), 0, s, L2, L1, nullptr, nullptr, nullptr, 0, 0ull);
        else if (S <= 256)
            hipLaunchKernelGGL((k_fill_scatter<true, 256, kFillTileWide>), dim3(grid_for(n, kFillTileWide, F1)),
                               dim3(kFillBlock), 0, s, L2, L1, nullptr, nullptr, nullptr, 0, 0ull);
        else
            hipLaunchKernelGGL((k_fill_scatter<true, kMaxSlicesPerRegion, kFillTile>),
                               dim3(grid_for(n, kFillTile, F1)), dim3(kFillBlock), 0, s, L2, L1, nullptr, nullptr,
                               nullptr, 0, 0ull);
        return check_launch("k_fill_scatter (level 2)");
    }
    hipLaunchKernelGGL((k_fill_scatter<false, kMaxFanoutWide, kFillTile>), dim3(grid_for(n, kFillTile, 0)),
                       dim3(kFillBlock), 0, s, L2, L2, keys, vals, pred, bound, n);
    return check_launch("k_fill_scatter");
}

// The fill-partitioned probe: no count passes; level 1 by fmix32(key) into
// fanout regions, level 2 by the same hash bits into the slices' regions,
// then the LDS probe per slice (slices of at most NB buckets) and the hashed
// probe of the overflow rows, both with accumulator Acc.
template <uint32_t NB, class Acc>
static int probe_lds_fill(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                          const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout, void *ws,
                          uint64_t ws_bytes, const LdsAccArgs &aa, mgv_launch LC, hipStream_t s) {
    for (const int32_t *c : {pred, keys, vals})
        if (c && !aligned16(c)) return fail(MGV_E_ALIGN, "probe_lds: column is not 16-byte aligned");
    FillLayout L1 = {}, L2 = {};
    uint32_t F2 = 0;
    uint64_t need = 0;
    int rc = fill_probe_plan(n, nbuckets, fanout, nullptr, 0, nullptr, nullptr, &F2, &need, NB);
    if (rc) return rc;
    if (!ws || ws_bytes < need) return fail(MGV_E_CAPACITY, "LDS probe: workspace too small");
    if ((rc = fill_probe_plan(n, nbuckets, fanout, ws, ws_bytes, &L1, &L2, &F2, nullptr, NB))) return rc;
    const uint64_t counters = (256 + (uint64_t)(fanout + F2) * 8 + 255) / 256 * 256;
    if ((rc = mgv_fill_u64((uint64_t *)ws, counters / 8, 0, (mgv_stream_t)s))) return rc;
    if ((rc = fill_partition(L1, L2, fanout, F2, keys, vals, pred, bound, n, s))) return rc;
    const uint32_t cus = device_cus();
    LdsProbeArgs a = {};
    a.slots = slots;
    a.nbuckets = nbuckets;
    a.slice = nbuckets / F2;
    a.nparts = F2;
    a.fill = L2.fill;
    a.cap = L2.cap;
    a.fcol[0] = L2.col[0];
    a.fcol[1] = L2.col[1];
    a.next = (uint32_t *)ws;
    if (!(LC.reserved & MGV_LAUNCH_KEY_SLICES) && tag_geometry(nbuckets, &a.geom)) {
        static thread_local int cached = 0;
        if (!cached) cached = persistent_blocks(k_lds_probe<SliceTag<NB>, Acc>, kLdsBlock);
        hipLaunchKernelGGL((k_lds_probe<SliceTag<NB>, Acc>), dim3(LC.grid ? LC.grid : (uint32_t)cached),
                           dim3(kLdsBlock), 0, s, a, aa);
    } else {
        static thread_local int cached = 0;
        if (!cached) cached = persistent_blocks(k_lds_probe<SliceT<NB>, Acc>, kLdsBlock);
        hipLaunchKernelGGL((k_lds_probe<SliceT<NB>, Acc>), dim3(LC.grid ? LC.grid : (uint32_t)cached),
                           dim3(kLdsBlock), 0, s, a, aa);
    }
    if ((rc = check_launch("k_lds_probe"))) return rc;
    // rows whose region was full at either level (none for a uniform hash
    // partition)
    const HashedRows o2 = {L2.col[0] + L2.ostart, L2.col[1] + L2.ostart, L2.ofill, L2.ocap};
    const HashedRows o1 = L1.F ? HashedRows{L1.col[0] + L1.ostart, L1.col[1] + L1.ostart, L1.ofill, L1.ocap}
                               : HashedRows{nullptr, nullptr, nullptr, 0};
    // one row per lane and a dependent bucket load per chain step: eight
    // blocks per CU keep enough loads in flight when skewed keys fill the
    // overflow areas (the blocks find no rows otherwise)
    hipLaunchKernelGGL(k_probe_hashed<Acc>, dim3(8 * cus), dim3(256), 0, s, slots, nbuckets, o2, o1, aa);
    return check_launch("k_probe_hashed");
}

// ------------------------------------------------------------------ LDS batch probe
//
// The probes of a batch whose tables split into slices of at most 2048
// buckets: every such query's selected rows are scattered (k_batch_scatter,
// hashed keys) into F regions, F = nbuckets / slice, region p holding the
// rows whose home bucket lies in slice p; one persistent launch then takes
// one (query, region) item per block iteration, copies the query's slice
// into LDS as key slices, probes the region's rows from it exactly as
// k_lds_probe does, and adds the counts to the query's result.  The probes
// never fetch a random line (a line miss costs about 1/55 G s on MI355X,
// the rate that bounds the fused batch's direct probes).

struct LdsBatchArgs {
    const mgv_query *Q;
    const mgv_work_item *items;   // partition items: .query, .lo = region index
    uint64_t nparts;
    const uint64_t *regions;      // kRegionWords per query: first row, rows per region, F, slice buckets
    const unsigned long long *fill;
    const int32_t *rhash, *rval;  // region columns: fmix32(key), val
    uint32_t *next;               // claim counter (zeroed before the launch)
};

__global__ __launch_bounds__(kLdsBlock) void k_batch_lds_probe(LdsBatchArgs a) {
    typedef SliceT<kSliceMax> Slice;
    __shared__ __align__(16) Slice S;
    __shared__ WaveQueue s_queue[kLdsBlock / kWave];
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    WaveQueue &Q = s_queue[t / kWave];
    const TagGeom geom = {0u, 0u, 0u};   // key slices: unused
    if (t == 0) S.pay[Slice::kZero] = 0u;   // visible after the first barrier below
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t it = s_part;
        __syncthreads();   // s_part is rewritten next round; the slice is free again
        if (it >= a.nparts) break;
        const mgv_work_item wi = a.items[it];
        const uint32_t qi = wi.query;
        const mgv_query q = a.Q[qi];
        const uint64_t *rg = a.regions + (uint64_t)kRegionWords * qi;
        const uint64_t cap = rg[1];
        const uint32_t slice = (uint32_t)rg[3];
        const uint32_t p = (uint32_t)wi.lo;
        const uint64_t r0 = rg[0] + (uint64_t)p * cap;
        const unsigned long long f = a.fill[(uint64_t)qi * kBatchMaxParts + p];
        const uint64_t r1 = r0 + (f < cap ? f : cap);
        if (r0 == r1) continue;
        const uint32_t nb = q.nbuckets;
        const uint32_t lo = p * slice;
        const uint32_t nload = min(slice + kSliceExtra, nb - lo);   // slice <= kSliceMax
        AccSum acc;
        v4i k[kAhead], v[kAhead];
        uint32_t m[kAhead];
        uint64_t i0 = r0 + 4ull * t;
#pragma unroll
        for (int x = 0; x < kAhead; ++x) {
            const uint64_t ix = i0 + (uint64_t)x * kQuadStride;
            m[x] = load_rows4(a.rhash, a.rval, ix < r1 ? ix : r0, ix < r1 ? r1 : r0, k[x], v[x]);
        }
        load_slice(S, q.slots, lo, nload, nb, geom);
        __syncthreads();
        // r0 is a multiple of 64 rows, so quads are 16-byte aligned
        uint32_t head = 0, tail = 0;   // the wave's queue (wave-uniform counters)
        while (ballot(i0 < r1)) {
            const uint64_t i1 = i0 + (uint64_t)kAhead * kQuadStride;
            v4i kn[kAhead], vn[kAhead];
            uint32_t mn[kAhead];
#pragma unroll
            for (int x = 0; x < kAhead; ++x) {
                const uint64_t ix = i1 + (uint64_t)x * kQuadStride;
                mn[x] = load_rows4(a.rhash, a.rval, ix < r1 ? ix : r0, ix < r1 ? r1 : r0, kn[x], vn[x]);
            }
#pragma unroll
            for (int x = 0; x < kAhead; ++x)
                lds_probe_quad(k[x], v[x], m[x], S, q.slots, nb, lo, nload, geom, acc, Q, head, tail);
#pragma unroll
            for (int x = 0; x < kAhead; ++x) {
                k[x] = kn[x];
                v[x] = vn[x];
                m[x] = mn[x];
            }
            i0 = i1;
        }
        while (tail != head) head = queue_round(Q, head, tail, S, q.slots, nb, lo, nload, geom, acc);
        acc.close(LdsAccArgs{q.res, nullptr, nullptr, 0});
        __syncthreads();
    }
}

// The builds of an LDS batch: one (query, slice) item per block iteration
// copies the slice (up to kBuildSlice buckets) into LDS, inserts the
// region's rows with LDS compare-and-swap in slot order (as k_lds_build), and
// writes the slice back; rows whose chain leaves the slice are appended to
// the overflow list, which k_batch_build_overflow inserts afterwards.  With
// fresh set every slice starts empty instead of being read, and every slice
// is written, rows or not: the tables need no clear before the call.
struct LdsBatchBuildArgs {
    const mgv_query *Q;
    const mgv_work_item *items;   // (query, slice) items of BUILD queries: .lo = slice
    uint64_t nparts;
    const uint64_t *regions;
    const unsigned long long *fill;
    const int32_t *rhash, *rval;
    uint32_t *next;
    unsigned long long *ocount;   // overflow list: {query, key, payload}
    uint32_t *oquery;
    int32_t *okey, *opay;
    uint32_t fresh;
};

__global__ __launch_bounds__(kLdsBlock) void k_batch_lds_build(LdsBatchBuildArgs a) {
    __shared__ __align__(16) unsigned long long tab[kBuildSlice * 8];
    __shared__ uint32_t s_part;
    const uint32_t t = threadIdx.x;
    for (;;) {
        if (t == 0) s_part = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t it = s_part;
        __syncthreads();
        if (it >= a.nparts) break;
        const mgv_work_item wi = a.items[it];
        const uint32_t qi = wi.query;
        const mgv_query q = a.Q[qi];
        const uint64_t *rg = a.regions + (uint64_t)kRegionWords * qi;
        const uint64_t cap = rg[1];
        const uint32_t slice = (uint32_t)rg[3];   // <= kBuildSlice
        const uint32_t p = (uint32_t)wi.lo;
        const uint64_t r0 = rg[0] + (uint64_t)p * cap;
        const unsigned long long f = a.fill[(uint64_t)qi * kBatchMaxParts + p];
        const uint64_t r1 = r0 + (f < cap ? f : cap);
        if (r0 == r1 && !a.fresh) continue;   // an untouched slice is not copied
        const uint32_t nb = q.nbuckets;
        const uint32_t lo = p * slice;
        v4u *g = reinterpret_cast<v4u *>(q.slots + (uint64_t)lo * 8);
        v4u *l = reinterpret_cast<v4u *>(tab);
        if (a.fresh) {
            const v4u e = {~0u, ~0u, ~0u, ~0u};   // kEmptySlot in every slot
            for (uint32_t i = t; i < slice * 4; i += kLdsBlock) l[i] = e;
        } else {
            for (uint32_t i = t; i < slice * 4; i += kLdsBlock) l[i] = load_global(g + i);
        }
        __syncthreads();
        uint64_t errs = 0;
        // r0 is a multiple of 64 rows; the region runs on to its capacity
        for (uint64_t i0 = r0 + 4ull * t; i0 < r1; i0 += 4ull * kLdsBlock) {
            const v4i h4 = load_stream4(a.rhash + i0);
            const v4i p4 = load_stream4(a.rval + i0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (i0 + j >= r1) break;
                const uint32_t h = (uint32_t)q_get(h4, j), pay = (uint32_t)q_get(p4, j);
                const uint32_t key = fmix32_inv(h);
                const uint64_t val = ((uint64_t)pay << 32) | key;
                if (val == kEmptySlot) {
                    ++errs;
                    continue;
                }
                const uint32_t idx = bucket_of_hash(h, nb) - lo;   // < slice: the region's rows
                if (idx >= slice || !lds_insert(tab, slice, idx, val)) {
                    const unsigned long long o = atomic_add_global(a.ocount, 1ull);
                    a.oquery[o] = qi;
                    a.okey[o] = (int32_t)key;
                    a.opay[o] = (int32_t)pay;
                }
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < slice * 4; i += kLdsBlock) g[i] = l[i];
        errs = wave_sum_u64(errs);
        if (lane_id() == 0 && errs && q.res)
            atomic_add_global((unsigned long long *)&q.res->errors, (unsigned long long)errs);
        __syncthreads();
    }
}

// The overflow list of an LDS batch build, inserted into each row's table
// after every slice is written back (chains continue into the next slices,
// or wrap to bucket 0).
__global__ __launch_bounds__(256) void k_batch_build_overflow(const mgv_query *Q, const unsigned long long *ocount,
                                                              const uint32_t *oquery, const int32_t *okey,
                                                              const int32_t *opay) {
    const uint64_t n = *ocount;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const mgv_query q = Q[oquery[i]];
        OpBuild<1> op;
        op.slots = q.slots;
        op.nbuckets = q.nbuckets;
        op.res = nullptr;
        op.errs = 0;
        op.insert((uint32_t)okey[i], (uint32_t)opay[i]);
        if (op.errs && q.res) atomic_add_global((unsigned long long *)&q.res->errors, (unsigned long long)op.errs);
    }
}

// Partitions of a query's table for the LDS batch: the smallest power of two
// F <= kBatchMaxParts that divides nbuckets into slices of at most smax
// buckets (kSliceMax for a probe, kBuildSlice for a build; 0: none).
static uint32_t lds_batch_parts(uint32_t nb, uint32_t smax) {
    for (uint32_t F = 1; F <= (uint32_t)kBatchMaxParts; F <<= 1)
        if (nb % F == 0 && nb / F <= smax && nb / F > 0) return F;
    return 0;
}

// Slices of the group-by probe: its LDS group table needs half the slice.
constexpr uint32_t kGroupSlice = 1024;
constexpr uint32_t kLdsMaxGroups = 4096;

// The probe of mgv_hashjoin_probe_lds on workspace ws (both entry points).
static int probe_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                     const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout, void *ws,
                     uint64_t ws_bytes, mgv_result *res, mgv_launch LC, mgv_stream_t st) {
    if (!slots || !keys || !res || nbuckets == 0) return fail(MGV_E_INVALID, "probe_lds: null argument");
    if (n == 0) return MGV_OK;
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "probe_lds: n must be < 2^31");
    if (!(LC.reserved & MGV_LAUNCH_COUNTED_PART)) {
        LdsAccArgs aa = {res, nullptr, nullptr, 0};
        return probe_lds_fill<kProbeSlice, AccSum>(slots, nbuckets, pred, bound, keys, vals, n, fanout, ws, ws_bytes,
                                                 aa, LC, stream_of(st));
    }
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
    if (!cached) cached = persistent_blocks(k_lds_probe<SliceT<kSliceMax>, AccSum>, kLdsBlock);
    const uint32_t grid = LC.grid ? LC.grid : (uint32_t)cached;
    LdsProbeArgs a = {};
    a.slots = slots;
    a.nbuckets = nbuckets;
    a.slice = nbuckets / F2;
    a.L = L2;
    a.nparts = L2.F;
    a.next = next;
    LdsAccArgs aa = {res, nullptr, nullptr, 0};
    hipLaunchKernelGGL((k_lds_probe<SliceT<kSliceMax>, AccSum>), dim3(grid), dim3(kLdsBlock), 0, s, a, aa);
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
    return probe_ws_need(n, nbuckets, fanout, bytes);
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
    uint64_t need = 0;
    int rc = probe_ws_need(n, nbuckets, fanout, &need);
    if (rc) return rc;
    if (need > kProbeScratchBytes)
        return fail(MGV_E_CAPACITY, "probe_lds_noalloc: %llu rows need %llu bytes of scratch, the array holds %llu",
                    (unsigned long long)n, (unsigned long long)need, (unsigned long long)kProbeScratchBytes);
    return probe_lds(slots, nbuckets, pred, bound, keys, vals, n, fanout, ds->probe_scratch, kProbeScratchBytes,
                     res, LC, st);
}

int mgv_probe_groupby_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes) {
    if (!bytes || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_probe_groupby_lds_workspace_size: bad argument");
    uint32_t F2 = 0;
    return fill_probe_plan(n, nbuckets, fanout, nullptr, 0, nullptr, nullptr, &F2, bytes, kGroupSlice);
}

int mgv_hashjoin_probe_groupby_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                                   uint32_t ngroups, uint32_t fanout, void *ws, uint64_t ws_bytes, mgv_launch LC,
                                   mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!slots || !keys || !vals || !gcnt || !gsum || nbuckets == 0 || ngroups == 0)
        return fail(MGV_E_INVALID, "probe_groupby_lds: bad argument");
    if (ngroups > kLdsMaxGroups) return fail(MGV_E_INVALID, "probe_groupby_lds: more than %u groups", kLdsMaxGroups);
    if (n == 0) return MGV_OK;
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "probe_groupby_lds: n must be < 2^31");
    LdsAccArgs aa = {nullptr, gcnt, gsum, ngroups};
    return probe_lds_fill<kGroupSlice, AccGroup<kLdsMaxGroups>>(slots, nbuckets, pred, bound, keys, vals, n, fanout,
                                                                ws, ws_bytes, aa, LC, stream_of(st));
}

int mgv_build_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes) {
    if (!bytes || nbuckets == 0) return fail(MGV_E_INVALID, "mgv_build_lds_workspace_size: bad argument");
    return build_ws_need(n, nbuckets, fanout, bytes);
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
    hipStream_t s = stream_of(st);
    const int32_t *pcol = payload_mode == MGV_PAYLOAD_COLUMN ? payload : nullptr;   // NONE: zeros
    for (const int32_t *c : {pred, keys, pcol})
        if (c && !aligned16(c)) return fail(MGV_E_ALIGN, "build_lds: column is not 16-byte aligned");
    uint32_t F2 = 0;
    int32_t *okey = nullptr, *opay = nullptr;
    int rc;
    if (!(LC.reserved & MGV_LAUNCH_COUNTED_PART)) {
        FillLayout F1L = {}, F2L = {};
        if ((rc = build_fill_plan(n, nbuckets, fanout, ws, ws_bytes, &F1L, &F2L, &F2, &okey, &opay, nullptr)))
            return rc;
        const uint64_t counters = (256 + (uint64_t)(fanout + F2) * 8 + 255) / 256 * 256;
        if ((rc = mgv_fill_u64((uint64_t *)ws, counters / 8, 0, (mgv_stream_t)s))) return rc;
        if ((rc = fill_partition(F1L, F2L, fanout, F2, keys, pcol, pred, bound, n, s))) return rc;
        LdsBuildArgs a = {};
        a.slots = slots;
        a.nbuckets = nbuckets;
        a.slice = nbuckets / F2;
        a.nparts = F2;
        a.fill = F2L.fill;
        a.cap = F2L.cap;
        a.fcol[0] = F2L.col[0];
        a.fcol[1] = F2L.col[1];
        a.next = (uint32_t *)ws;
        a.ovf_count = (unsigned long long *)ws + 3;
        a.ovf_key = okey;
        a.ovf_pay = opay;
        a.res = res;
        if (a.slice <= kBuildSlice) {
            static thread_local int cached = 0;
            if (!cached) cached = persistent_blocks(k_lds_build<kBuildSlice>, kLdsBlock);
            hipLaunchKernelGGL(k_lds_build<kBuildSlice>, dim3(LC.grid ? LC.grid : (uint32_t)cached),
                               dim3(kLdsBlock), 0, s, a);
        } else {
            static thread_local int cached = 0;
            if (!cached) cached = persistent_blocks(k_lds_build<kSliceMax>, kLdsBlock);
            hipLaunchKernelGGL(k_lds_build<kSliceMax>, dim3(LC.grid ? LC.grid : (uint32_t)cached), dim3(kLdsBlock),
                               0, s, a);
        }
        if ((rc = check_launch("k_lds_build"))) return rc;
        // rows whose region was full at either level (duplicate or skewed
        // keys), then rows whose chain left their slice
        const HashedRows o2 = {F2L.col[0] + F2L.ostart, F2L.col[1] + F2L.ostart, F2L.ofill, F2L.ocap};
        const HashedRows o1 = F1L.F ? HashedRows{F1L.col[0] + F1L.ostart, F1L.col[1] + F1L.ostart, F1L.ofill,
                                                 F1L.ocap}
                                    : HashedRows{nullptr, nullptr, nullptr, 0};
        hipLaunchKernelGGL(k_lds_build_overflow, dim3(4 * device_cus()), dim3(256), 0, s, a, o2, o1);
        return check_launch("k_lds_build_overflow");
    }
    PartLayout L1, L2;
    if ((rc = build_plan(n, nbuckets, fanout, ws, ws_bytes, &L1, &L2, &F2, &okey, &opay, nullptr))) return rc;
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
    LdsBuildArgs a = {};
    a.slots = slots;
    a.nbuckets = nbuckets;
    a.slice = nbuckets / F2;
    a.L = L2;
    a.nparts = L2.F;
    a.next = ctr;
    a.ovf_count = (unsigned long long *)(ctr + 2);
    a.ovf_key = okey;
    a.ovf_pay = opay;
    a.res = res;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(k_lds_build<kSliceMax>, kLdsBlock);
    hipLaunchKernelGGL(k_lds_build<kSliceMax>, dim3(LC.grid ? LC.grid : (uint32_t)cached), dim3(kLdsBlock), 0, s, a);
    if ((rc = check_launch("k_lds_build"))) return rc;
    const HashedRows none = {nullptr, nullptr, nullptr, 0};
    hipLaunchKernelGGL(k_lds_build_overflow, dim3(256), dim3(256), 0, s, a, none, none);
    return check_launch("k_lds_build_overflow");
}

int mgv_batch_plan_lds(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, uint64_t min_rows,
                       mgv_work_item *items, uint64_t cap, uint64_t *regions, mgv_lds_batch_plan *plan) {
    if (!queries || !plan) return fail(MGV_E_INVALID, "mgv_batch_plan_lds: null argument");
    uint64_t n = 0;
    int rc = mgv_batch_plan(queries, nqueries, vector_rows, nullptr, 0, &n);   // validates the queries
    if (rc) return rc;
    auto parts = [&](const mgv_query &q) -> uint32_t {
        if (q.n == 0 || q.n < min_rows) return 0;
        if (q.kind == MGV_Q_PROBE_SUM && !(q.flags & MGV_PROBE_DIGEST)) return lds_batch_parts(q.nbuckets, kSliceMax);
        if (q.kind == MGV_Q_BUILD && q.payload_mode != MGV_PAYLOAD_ROWID)
            return lds_batch_parts(q.nbuckets, kBuildSlice);
        return 0;
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
    for (uint32_t qi = 0; qi < nqueries; ++qi) {
        const uint32_t F = parts(queries[qi]);
        const uint64_t c = F ? batch_region_cap(queries[qi].n, F) : 0;
        if (regions) {
            regions[kRegionWords * qi] = rows;
            regions[kRegionWords * qi + 1] = c;
            regions[kRegionWords * qi + 2] = F;
            regions[kRegionWords * qi + 3] = F ? queries[qi].nbuckets / F : 0;
        }
        rows += (uint64_t)F * c;
    }
    for (uint32_t qi = 0; qi < nqueries; ++qi)
        if (parts(queries[qi]))
            for (uint64_t lo = 0; lo < queries[qi].n; lo += kBatchTileWide)
                put(qi, lo, queries[qi].n - lo < (uint64_t)kBatchTileWide ? queries[qi].n - lo : kBatchTileWide);
    plan->nscatter = k;
    uint64_t build_rows = 0;
    for (uint32_t qi = 0; qi < nqueries; ++qi) {
        const uint32_t F = parts(queries[qi]);
        if (queries[qi].kind != MGV_Q_BUILD)
            for (uint32_t p = 0; p < F; ++p) put(qi, p, 0);
    }
    plan->nparts = k - plan->nscatter;
    for (uint32_t qi = 0; qi < nqueries; ++qi) {
        const uint32_t F = parts(queries[qi]);
        if (queries[qi].kind == MGV_Q_BUILD && F) {
            for (uint32_t p = 0; p < F; ++p) put(qi, p, 0);
            build_rows += queries[qi].n;
        }
    }
    plan->nbuild_parts = k - plan->nscatter - plan->nparts;
    plan->build_rows = build_rows;
    for (uint32_t qi = 0; qi < nqueries; ++qi)
        if (!parts(queries[qi]))
            for (uint64_t lo = 0; lo < queries[qi].n; lo += vector_rows)
                put(qi, lo, queries[qi].n - lo < vector_rows ? queries[qi].n - lo : vector_rows);
    plan->nitems = k;
    plan->nqueries = nqueries;
    plan->reserved = 0;
    plan->region_rows = rows;
    plan->workspace_bytes = batch_ws_bytes(nqueries, rows) + lds_batch_ovf_bytes(build_rows);
    if (items && k > cap) return fail(MGV_E_CAPACITY, "mgv_batch_plan_lds: %llu items > capacity %llu",
                                      (unsigned long long)k, (unsigned long long)cap);
    return MGV_OK;
}

int mgv_batch_run_lds(const mgv_query *queries, const mgv_work_item *items, const uint64_t *regions,
                      const mgv_lds_batch_plan *plan, void *ws, uint64_t ws_bytes, mgv_launch L, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!queries || !plan || !regions || !ws || (plan->nitems && !items))
        return fail(MGV_E_INVALID, "mgv_batch_run_lds: null argument");
    if (plan->nitems >= 0xffffffffull || plan->nscatter + plan->nparts + plan->nbuild_parts > plan->nitems)
        return fail(MGV_E_INVALID, "mgv_batch_run_lds: inconsistent plan");
    const uint64_t need = batch_ws_bytes(plan->nqueries, plan->region_rows) + lds_batch_ovf_bytes(plan->build_rows);
    if (ws_bytes < need) return fail(MGV_E_CAPACITY, "mgv_batch_run_lds: workspace %llu < %llu bytes",
                                     (unsigned long long)ws_bytes, (unsigned long long)need);
    if (!aligned16(ws)) return fail(MGV_E_ALIGN, "mgv_batch_run_lds: workspace not 16-byte aligned");
    if (plan->nitems == 0) return MGV_OK;
    hipStream_t s = stream_of(st);
    char *w = (char *)ws;
    const uint64_t fills = batch_fill_bytes(plan->nqueries);
    const uint64_t col = (plan->region_rows * 4 + 255) / 256 * 256;
    // counters: [0] scatter claims, [1] probe partition claims, [2] build
    // partition claims, 64-bit word 2 the overflow count, [8] the direct launch
    uint32_t *ctr = (uint32_t *)w;
    unsigned long long *ocount = (unsigned long long *)w + 2;
    unsigned long long *fill = (unsigned long long *)(w + 256);
    int32_t *rhash = (int32_t *)(w + 256 + fills), *rval = (int32_t *)(w + 256 + fills + col);
    // the overflow list of the builds: query indices, keys, payloads
    const uint64_t ob = (plan->build_rows * 4 + 255) / 256 * 256;
    char *ov = w + batch_ws_bytes(plan->nqueries, plan->region_rows);
    uint32_t *oquery = (uint32_t *)ov;
    int32_t *okey = (int32_t *)(ov + ob), *opay = (int32_t *)(ov + 2 * ob);
    // counters and fill counts: a fill kernel (the call can be captured into a graph)
    int rc = mgv_fill_u64((uint64_t *)ws, (256 + fills) / 8, 0, (mgv_stream_t)s);
    if (rc) return rc;
    const BatchScatter S = {regions, fill, rhash, rval, ctr, plan->nscatter, 1u, (uint32_t)kBatchTileWide,
                            ocount, oquery, okey, opay};
    if ((rc = launch_batch_scatter(queries, items, S, s))) return rc;
    if (plan->nparts) {
        LdsBatchArgs a = {queries, items + plan->nscatter, plan->nparts, regions, fill, rhash, rval, ctr + 1};
        static thread_local int cached = 0;
        if (!cached) cached = persistent_blocks(k_batch_lds_probe, kLdsBlock);
        const uint64_t grid = plan->nparts < (uint64_t)cached ? plan->nparts : (uint64_t)cached;
        hipLaunchKernelGGL(k_batch_lds_probe, dim3((uint32_t)grid), dim3(kLdsBlock), 0, s, a);
        if ((rc = check_launch("k_batch_lds_probe"))) return rc;
    }
    if (plan->nbuild_parts) {
        LdsBatchBuildArgs b = {queries, items + plan->nscatter + plan->nparts, plan->nbuild_parts, regions, fill,
                               rhash, rval, ctr + 2, ocount, oquery, okey, opay,
                               (L.reserved & MGV_LAUNCH_FRESH_TABLES) ? 1u : 0u};
        static thread_local int cached = 0;
        if (!cached) cached = persistent_blocks(k_batch_lds_build, kLdsBlock);
        const uint64_t grid = plan->nbuild_parts < (uint64_t)cached ? plan->nbuild_parts : (uint64_t)cached;
        hipLaunchKernelGGL(k_batch_lds_build, dim3((uint32_t)grid), dim3(kLdsBlock), 0, s, b);
        if ((rc = check_launch("k_batch_lds_build"))) return rc;
        hipLaunchKernelGGL(k_batch_build_overflow, dim3(4 * device_cus()), dim3(256), 0, s, queries, ocount, oquery,
                           okey, opay);
        if ((rc = check_launch("k_batch_build_overflow"))) return rc;
    }
    const uint64_t done = plan->nscatter + plan->nparts + plan->nbuild_parts;
    const uint64_t nd = plan->nitems - done;
    if (!nd) return MGV_OK;
    return mgv_batch_run(queries, items + done, nd, ctr + 8, L, st);
}

}  // extern "C"
