// mgv_partition.hip -- radix-partitioned, XCD-aware probe.
//
// A table larger than the L2 makes every probe a 128-byte line fetch from
// HBM.  This path first scatters the probe rows into F partitions by the top
// bits of fmix32(key) -- the same bits that pick the bucket, so partition p
// only touches the p-th 1/F slice of the bucket array -- and then probes
// partition by partition, each XCD draining the partitions p = xcd (mod 8) so
// that an XCD's active table slices stay in its own L2.  The optional select
// (pred <= bound) is applied while partitioning, so only survivors move.
//
// Kernels (one stream, in order): k_part_prep (zero counters),
// k_part_count (per-block histograms), k_part_scan (aligned partition
// starts), k_part_scatter (LDS-staged scatter), k_part_probe (persistent
// work-queue probe).  The reference's counterpart of the partitioning step is
// split's ballot-compacted per-destination staging (src/operators/split.cu:
// 54-118), there used for 2-GPU repartitioning with hash3(x) = x & 1.
#include <hip/hip_runtime.h>

#include "mgv_internal.hpp"
#include "mgv_part.hpp"
#include "mgv_pipes.hpp"

namespace mgv {

__device__ __forceinline__ uint32_t part_of(int32_t key, uint32_t F) {
    return (uint32_t)(((uint64_t)fmix32((uint32_t)key) * F) >> 32);
}

// Partition of a key: hash partitioning (joins) or key ranges (group-by:
// partition = key >> rshift, keys outside [0, rlimit) are dropped).
__device__ __forceinline__ uint32_t part_key(const PartIn &in, int32_t key, uint32_t F) {
    if (in.rlimit) return (uint32_t)key >> in.rshift;
    return in.key_hash == 2 ? part_of_hash(key, F) : part_of(key, F);
}

// Bit j set when row j of a quad has a key the partitioning accepts.
__device__ __forceinline__ uint32_t key_mask(const PartIn &in, const v4i &k) {
    if (!in.rlimit) return 0xfu;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)((uint32_t)q_get(k, j) < in.rlimit) << j;
    return m;
}



__host__ __device__ static inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

int plan_layout(uint64_t n, uint32_t F, uint32_t ncols, void *ws, uint64_t ws_bytes, PartLayout *L,
                uint64_t *need, uint32_t align) {
    if (F < 2 || F > kMaxFanoutWide || (F & (F - 1)))
        return fail(MGV_E_INVALID, "fan-out %u not a power of two in [2,%u]", F, kMaxFanoutWide);
    if (align < 1 || align > 64 || (align & (align - 1))) return fail(MGV_E_INVALID, "partition alignment %u", align);
    if (ncols < 1 || ncols > 3) return fail(MGV_E_INVALID, "partitioned columns %u not in [1,3]", ncols);
    uint64_t nblk = (n + 16383) / 16384;
    if (nblk < 1) nblk = 1;
    if (nblk > kMaxPartBlocks) nblk = kMaxPartBlocks;
    uint64_t rpb = align_up((n + nblk - 1) / nblk, kTileMax);
    nblk = (n + rpb - 1) / rpb;
    if (nblk < 1) nblk = 1;
    const uint64_t cap = align_up(n + (uint64_t)F * align, 64);
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint64_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const uint64_t o_hist = take(nblk * F * 4);
    const uint64_t o_ptot = take(F * 8);
    const uint64_t o_pst = take((F + 1) * 8);
    const uint64_t o_pcur = take(F * 8);
    const uint64_t o_xh = take(64);
    uint64_t o_col[3] = {0, 0, 0};
    for (uint32_t c = 0; c < ncols; ++c) o_col[c] = take(cap * 4);
    if (need) *need = off;
    if (!L) return MGV_OK;
    if (!ws || ws_bytes < off)
        return fail(MGV_E_CAPACITY, "partition workspace %llu bytes < %llu needed", (unsigned long long)ws_bytes,
                    (unsigned long long)off);
    char *b = (char *)ws;
    L->F = F;
    L->align = align;
    L->nblk = (uint32_t)nblk;
    L->ncols = ncols;
    L->rows_per_blk = rpb;
    L->n = n;
    L->cap = cap;
    L->hist = (uint32_t *)(b + o_hist);
    L->ptotal = (unsigned long long *)(b + o_ptot);
    L->pstart = (unsigned long long *)(b + o_pst);
    L->pcursor = (unsigned long long *)(b + o_pcur);
    L->xhead = (uint32_t *)(b + o_xh);
    for (int c = 0; c < 3; ++c) L->col[c] = c < (int)ncols ? (int32_t *)(b + o_col[c]) : nullptr;
    return MGV_OK;
}



__global__ void k_part_prep(PartLayout L) {
    for (uint32_t p = threadIdx.x; p < L.F; p += blockDim.x) L.ptotal[p] = 0;
    if (threadIdx.x < 8) L.xhead[threadIdx.x] = 0;
}

__device__ __forceinline__ uint32_t quad_mask(const PartIn &in, uint64_t i0, uint64_t hi, v4i *predq) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)(i0 + j < hi) << j;
    if (in.pred) {
        v4i c = load_quad(in.pred, i0, hi);
        *predq = c;
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) s |= (uint32_t)(q_get(c, j) <= in.bound) << j;
        m &= s;
    }
    return m;
}

// Per-block partition histograms.  The block's rows are read in steps of
// 4 * kPartBlock rows: whole steps with unconditional quad loads, the next
// step's loads issued before the current step is counted, then the last,
// partial step with bounds-checked loads.
template <uint32_t MAXF>
__global__ __launch_bounds__(kPartBlock) void k_part_count(PartLayout L, PartIn in) {
    __shared__ uint32_t h[MAXF];
    __shared__ uint32_t s_err;
    for (uint32_t p = threadIdx.x; p < L.F; p += kPartBlock) h[p] = 0;
    if (threadIdx.x == 0) s_err = 0;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * L.rows_per_blk;
    uint64_t hi = min(L.n, lo + L.rows_per_blk);
    if (in.nlimit) hi = min(hi, (uint64_t)*in.nlimit);
    uint32_t errs = 0;
    auto count_quad = [&](const v4i &k, uint32_t m) {
        const uint32_t ok = m & key_mask(in, k);
        errs += (uint32_t)__builtin_popcount(m & ~ok);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if ((ok >> j) & 1u) atomicAdd(&h[part_key(in, q_get(k, j), L.F)], 1u);
    };
    auto pred_mask = [&](const v4i &c) {
        uint32_t m = 0xfu;
        if (in.pred) {
#pragma unroll
            for (int j = 0; j < 4; ++j) m &= ~((uint32_t)(q_get(c, j) > in.bound) << j);
        }
        return m;
    };
    constexpr uint64_t kStep = 4ull * kPartBlock;
    const uint64_t whole = hi > lo ? (hi - lo) / kStep : 0;
    uint64_t i0 = lo + 4ull * threadIdx.x;
    if (whole) {
        v4i k = load_stream4(in.key + i0);
        v4i c = in.pred ? load_stream4(in.pred + i0) : v4i{0, 0, 0, 0};
        for (uint64_t j = 0; j < whole; ++j) {
            // the last whole step reloads itself (in bounds, unused) so the
            // prefetch is unconditional
            const uint64_t i1 = j + 1 < whole ? i0 + kStep : i0;
            const v4i kn = load_stream4(in.key + i1);
            const v4i cn = in.pred ? load_stream4(in.pred + i1) : v4i{0, 0, 0, 0};
            count_quad(k, pred_mask(c));
            k = kn;
            c = cn;
            i0 += kStep;
        }
    }
    if (i0 < hi) {
        v4i c;
        const uint32_t m = quad_mask(in, i0, hi, &c);
        if (m) count_quad(load_quad(in.key, i0, hi), m);
    }
    if (errs) atomicAdd(&s_err, errs);
    __syncthreads();
    if (threadIdx.x == 0 && s_err && in.res)
        atomic_add_global((unsigned long long *)&in.res->errors, (unsigned long long)s_err);
    for (uint32_t p = threadIdx.x; p < L.F; p += kPartBlock) {
        uint32_t c = h[p];
        L.hist[(uint64_t)blockIdx.x * L.F + p] = c;
        if (c) atomic_add_global(&L.ptotal[p], (unsigned long long)c);
    }
}

// One block of kScanThreads threads: exclusive scan of the partition totals,
// each rounded up to the alignment; thread t owns E = F / kScanThreads
// consecutive partitions (or one, for F <= kScanThreads).
constexpr uint32_t kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_part_scan(PartLayout L) {
    __shared__ unsigned long long s[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t E = L.F > kScanThreads ? L.F / kScanThreads : 1;
    unsigned long long mine = 0;
    for (uint32_t e = 0; e < E; ++e) {
        const uint32_t p = t * E + e;
        if (p < L.F) mine += align_up(L.ptotal[p], L.align);
    }
    s[t] = mine;
    __syncthreads();
    for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
        const unsigned long long a = (t >= d) ? s[t - d] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    unsigned long long start = s[t] - mine;   // exclusive
    for (uint32_t e = 0; e < E; ++e) {
        const uint32_t p = t * E + e;
        if (p < L.F) {
            L.pstart[p] = start;
            L.pcursor[p] = start;
            start += align_up(L.ptotal[p], L.align);
        }
    }
    if (t == kScanThreads - 1) L.pstart[L.F] = s[t];
}

// LDS-staged scatter: each tile of TILE rows is counting-sorted by partition
// in LDS, then written out so that consecutive lanes write consecutive slots
// of a partition's run.
//
// Per tile: (1) rank every row within its partition with an LDS atomic;
// (2) the thread owning partitions [t*per, (t+1)*per) scans the tile counts,
// publishes base[p] = (block's next output row of p) - (p's first staged
// slot), and advances its cursors, kept in registers; (3) stage the rows in
// partition order; (4) write staged slot i to base[p] + i.
//
// Overlap: the next tile's columns are loaded before the current tile is
// ranked, and nothing reads them until the next iteration (masks, partition
// ids and row ids are computed at consume time).  The write step issues a
// fixed number of stores per thread (lanes past the tile's last staged row
// repeat that row's store), so the compiler can wait for the prefetched
// loads without also waiting for the stores, which share the same counter.
template <bool ROWID, uint32_t MAXF, int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void k_part_scatter(PartLayout L, PartIn in) {
    constexpr int Q = TILE / (4 * BLOCK);   // quads per thread per tile
    constexpr int R = TILE / BLOCK;         // write rounds per tile
    constexpr int PER = MAXF > (uint32_t)BLOCK ? (int)(MAXF / BLOCK) : 1;   // partitions per owner thread
    static_assert(Q >= 1 && TILE % (4 * BLOCK) == 0 && kTileMax % TILE == 0, "scatter shape");
    static_assert(TILE <= 65536, "tile offsets are 16-bit");
    __shared__ unsigned long long base[MAXF];
    __shared__ uint32_t tcnt[MAXF];
    __shared__ uint16_t toff[MAXF];   // first staged slot of each partition (< TILE)
    __shared__ uint32_t wtot[BLOCK / kWave];
    __shared__ uint32_t s_total;
    __shared__ int32_t sk[TILE];
    __shared__ int32_t sv[TILE];
    __shared__ int32_t sr[ROWID ? TILE : 1];
    const uint32_t F = L.F;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (F + BLOCK - 1) / BLOCK;   // <= PER
    unsigned long long cur[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        cur[e] = 0;
        const uint32_t p = t * per + e;
        if (e < (int)per && p < F) {
            const uint32_t c = L.hist[(uint64_t)blockIdx.x * F + p];
            cur[e] = c ? atomic_add_global(&L.pcursor[p], (unsigned long long)c) : 0;
            tcnt[p] = 0;
        }
    }
    const uint64_t lo = (uint64_t)blockIdx.x * L.rows_per_blk;
    uint64_t hi = min(L.n, lo + L.rows_per_blk);
    if (in.nlimit) hi = min(hi, (uint64_t)*in.nlimit);
    v4i nk[Q], nv[Q], nc[Q];
    uint32_t nb[Q];   // in-bounds rows of each quad
    auto load_tile = [&](uint64_t tb) {
        if (tb + TILE <= hi) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const uint64_t i0 = tb + 4ull * t + (uint64_t)q * 4 * BLOCK;
                nb[q] = 0xfu;
                nk[q] = load_stream4(in.key + i0);
                nv[q] = in.val ? load_stream4(in.val + i0) : v4i{0, 0, 0, 0};
                nc[q] = in.pred ? load_stream4(in.pred + i0) : v4i{0, 0, 0, 0};
            }
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const uint64_t i0 = tb + 4ull * t + (uint64_t)q * 4 * BLOCK;
                uint32_t m = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) m |= (uint32_t)(i0 + j < hi) << j;
                nb[q] = m;
                nk[q] = load_quad(in.key, i0, hi);
                nv[q] = in.val ? load_quad(in.val, i0, hi) : v4i{0, 0, 0, 0};
                nc[q] = in.pred ? load_quad(in.pred, i0, hi) : v4i{0, 0, 0, 0};
            }
        }
    };
    if (lo < hi) load_tile(lo);
    // The first tile is waited for here: entering the loop with its loads
    // pending would make the compiler wait for the previous iteration's
    // stores as well at the top of every iteration.
#pragma unroll
    for (int q = 0; q < Q; ++q)
        asm volatile("" ::"v"(nk[q].x), "v"(nk[q].y), "v"(nk[q].z), "v"(nk[q].w), "v"(nv[q].x), "v"(nv[q].y),
                     "v"(nv[q].z), "v"(nv[q].w), "v"(nc[q].x), "v"(nc[q].y), "v"(nc[q].z), "v"(nc[q].w));
    __syncthreads();
    for (uint64_t tb = lo; tb < hi; tb += TILE) {
        v4i ck[Q], cv[Q];
        uint32_t cm[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            ck[q] = nk[q];
            cv[q] = nv[q];
            uint32_t m = nb[q];
            if (in.pred) {
#pragma unroll
                for (int j = 0; j < 4; ++j) m &= ~((uint32_t)(q_get(nc[q], j) > in.bound) << j);
            }
            cm[q] = m & key_mask(in, ck[q]);
        }
        if (tb + TILE < hi) load_tile(tb + TILE);
        // 1. rank every row within its partition
        // masked rows get partition 0 (a dropped key may map past F); the
        // partition is recomputed from the key when the row is staged
        uint32_t rank[Q][4];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (cm[q] >> j) & 1u;
                rank[q][j] = ok ? atomicAdd(&tcnt[ok ? part_key(in, q_get(ck[q], j), F) : 0u], 1u) : 0u;
            }
        __syncthreads();
        // 2. owners scan the tile counts, publish toff / base, advance cursors
        uint32_t c[PER];
        uint32_t mine = 0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t p = t * per + e;
            c[e] = (e < (int)per && p < F) ? tcnt[p] : 0u;
            mine += c[e];
        }
        uint32_t start = scan_block_excl<BLOCK>(mine, wtot);
        if (t == BLOCK - 1) s_total = start + mine;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t p = t * per + e;
            if (e < (int)per && p < F) {
                toff[p] = (uint16_t)start;
                base[p] = cur[e] - start;
                cur[e] += c[e];
                tcnt[p] = 0;
                start += c[e];
            }
        }
        __syncthreads();
        // 3. stage rows in partition order (all slot reads first, then writes)
        uint32_t pos[Q][4];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (cm[q] >> j) & 1u;
                pos[q][j] = toff[ok ? part_key(in, q_get(ck[q], j), F) : 0u] + rank[q][j];
            }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((cm[q] >> j) & 1u) {
                    const int32_t key = q_get(ck[q], j);
                    sk[pos[q][j]] = in.key_hash == 1 ? (int32_t)fmix32((uint32_t)key) : key;
                    sv[pos[q][j]] = q_get(cv[q], j);
                    if (ROWID)
                        sr[pos[q][j]] = (int32_t)(in.row_base +
                                                  (int64_t)(tb + 4ull * t + (uint64_t)q * 4 * BLOCK + j));
                }
            }
        }
        __syncthreads();
        // 4. write the runs: exactly R stores per column and thread.  An
        // empty tile stores to slot cap - 1, which no partition reaches
        // (partitions end at pstart[F] <= n + F * (align - 1) < cap - 1); a
        // keys-only layout stores the key twice instead of a second column.
        const uint32_t total = s_total;
        int32_t *const c1 = L.ncols > 1 ? L.col[1] : L.col[0];
        constexpr int G = R < 8 ? R : 8;   // rounds per group: bounds the registers held
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += G) {
            int32_t k[G], v[G], rid[G];
            uint32_t ii[G];
#pragma unroll
            for (int r = 0; r < G; ++r) {
                ii[r] = total ? min(t + (uint32_t)(r0 + r) * BLOCK, total - 1) : 0u;
                k[r] = sk[ii[r]];
                v[r] = sv[ii[r]];
                rid[r] = ROWID ? sr[ii[r]] : 0;
            }
            unsigned long long dst[G];
#pragma unroll
            for (int r = 0; r < G; ++r) {
                // the staged value is the hash when key_hash is set
                const uint32_t p = total ? (in.key_hash ? part_of_hash(k[r], F) : part_key(in, k[r], F)) : 0u;
                dst[r] = total ? base[p] + ii[r] : L.cap - 1;
            }
#pragma unroll
            for (int r = 0; r < G; ++r) {
                L.col[0][dst[r]] = k[r];
                c1[dst[r]] = L.ncols > 1 ? v[r] : k[r];
                if (ROWID) L.col[2][dst[r]] = rid[r];
            }
        }
        __syncthreads();
    }
}

// Persistent probe over the partitions.  Each block copies the partition
// table into LDS and builds, for every XCD x, the cumulative item counts of
// its partitions p = x, x + 8, ...; afterwards every wave claims work items
// (kItemRows rows of one partition) on its own with one atomic, first from
// its own XCD's queue, then from the others'.
template <class Pipe, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_part_probe(typename Pipe::Params prm, PartLayout L) {
    constexpr int WAVES = BLOCK / kWave;
    typedef typename Pipe::template Op<WAVES> Op;
    __shared__ __align__(16) typename Op::Smem sm;
    __shared__ unsigned long long s_start[kMaxFanout];
    __shared__ unsigned long long s_size[kMaxFanout];
    __shared__ uint32_t s_cum[kMaxFanout];   // cumulative items, in per-XCD order
    __shared__ uint32_t s_tot[8];
    Op op;
    Pipe::setup(op, prm, &sm);
    const uint32_t F = L.F;
    for (uint32_t p = threadIdx.x; p < F; p += BLOCK) {
        s_start[p] = L.pstart[p];
        s_size[p] = L.ptotal[p];
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const uint32_t x = threadIdx.x;
        uint32_t acc = 0;
        for (uint32_t p = x; p < F; p += 8) {
            acc += (uint32_t)((s_size[p] + kItemRows - 1) / kItemRows);
            s_cum[p] = acc;
        }
        s_tot[x] = acc;
    }
    op.consume_open();   // contains a block barrier
    __syncthreads();
    const uint32_t home = xcc_id();
    Cols cols = {{L.col[0], L.ncols > 1 ? L.col[1] : nullptr, L.ncols > 2 ? L.col[2] : nullptr, nullptr}, -1};
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t x = (home + k) & 7u;
        const uint32_t tot = s_tot[x];
        const uint32_t np = F > x ? (F - x + 7) / 8 : 0;
        for (;;) {
            uint32_t item = 0;
            if (lane_id() == 0) item = atomicAdd(&L.xhead[x], 1u);
            item = __builtin_amdgcn_readlane(item, 0);
            if (item >= tot) break;
            // smallest m with s_cum[x + 8m] > item
            uint32_t a = 0, b = np - 1;
            while (a < b) {
                const uint32_t mid = (a + b) / 2;
                if (s_cum[x + 8 * mid] > item) b = mid; else a = mid + 1;
            }
            const uint32_t p = x + 8 * a;
            const uint32_t first = a ? s_cum[x + 8 * (a - 1)] : 0;
            const uint64_t ilo = s_start[p] + (uint64_t)(item - first) * kItemRows;
            const uint64_t pend = s_start[p] + s_size[p];
            const uint64_t ihi = (ilo + kItemRows < pend) ? ilo + kItemRows : pend;
            drive_chunks<Pipe::kCols>(op, cols, ilo, kChunk, ihi, 0);
        }
    }
    op.consume_close();
}

}  // namespace mgv

// ------------------------------------------------------------------ host side

namespace mgv {

template <class Pipe, int B>
static int part_probe_b(const typename Pipe::Params &prm, const PartLayout &L, const mgv_launch &LC, hipStream_t s) {
    auto kern = k_part_probe<Pipe, B>;
    static thread_local int cached = 0;
    if (!cached) cached = persistent_blocks(kern, B);
    const uint32_t grid = LC.grid ? LC.grid : (uint32_t)cached;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(B), 0, s, prm, L);
    return check_launch("k_part_probe");
}

template <class Pipe>
static int part_probe(const typename Pipe::Params &prm, const PartLayout &L, const mgv_launch &LC, hipStream_t s) {
    switch (LC.block) {
        case 64: return part_probe_b<Pipe, 64>(prm, L, LC, s);
        case 0:
        case 256: return part_probe_b<Pipe, 256>(prm, L, LC, s);
        case 1024: return part_probe_b<Pipe, 1024>(prm, L, LC, s);
        default: return fail(MGV_E_INVALID, "block %u not in {64,256,1024}", LC.block);
    }
}

template <bool ROWID, uint32_t MAXF>
static void launch_scatter(const PartLayout &L, const PartIn &in, uint32_t shape, hipStream_t s) {
    switch (shape) {
        case MGV_SCATTER_512x8192:
            hipLaunchKernelGGL((k_part_scatter<ROWID, MAXF, 512, 8192>), dim3(L.nblk), dim3(512), 0, s, L, in);
            break;
        case MGV_SCATTER_1024x8192:
            hipLaunchKernelGGL((k_part_scatter<ROWID, MAXF, 1024, 8192>), dim3(L.nblk), dim3(1024), 0, s, L, in);
            break;
        case MGV_SCATTER_1024x4096:
            hipLaunchKernelGGL((k_part_scatter<ROWID, MAXF, 1024, 4096>), dim3(L.nblk), dim3(1024), 0, s, L, in);
            break;
        default:
            hipLaunchKernelGGL((k_part_scatter<ROWID, MAXF, 256, 4096>), dim3(L.nblk), dim3(256), 0, s, L, in);
    }
}

template <uint32_t MAXF>
static int partition_rows_f(const PartLayout &L, const PartIn &in, uint32_t shape, hipStream_t s) {
    hipLaunchKernelGGL(k_part_prep, dim3(1), dim3(1024), 0, s, L);
    int rc = check_launch("k_part_prep");
    if (rc) return rc;
    hipLaunchKernelGGL(k_part_count<MAXF>, dim3(L.nblk), dim3(kPartBlock), 0, s, L, in);
    if ((rc = check_launch("k_part_count"))) return rc;
    hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(kScanThreads), 0, s, L);
    if ((rc = check_launch("k_part_scan"))) return rc;
    if (in.want_rowid) launch_scatter<true, MAXF>(L, in, 0, s);   // row ids: default shape only
    else launch_scatter<false, MAXF>(L, in, shape, s);
    return check_launch("k_part_scatter");
}

int partition_rows(const PartLayout &L, const PartIn &in, hipStream_t s, uint32_t shape) {
    for (const int32_t *p : {in.pred, in.key, in.val})
        if (p && !aligned16(p)) return fail(MGV_E_ALIGN, "partitioned column is not 16-byte aligned");
    if (L.F <= kMaxFanout) return partition_rows_f<kMaxFanout>(L, in, shape, s);
    if (in.want_rowid) return fail(MGV_E_INVALID, "row ids are partitioned with fan-out <= %u only", kMaxFanout);
    return partition_rows_f<kMaxFanoutWide>(L, in, shape, s);
}

__global__ void k_xcc_histogram(uint32_t *out) {
    if (threadIdx.x == 0) atomicAdd(&out[xcc_id()], 1u);
}

}  // namespace mgv

using namespace mgv;

extern "C" {

int mgv_partition_workspace_size(uint64_t n, uint32_t fanout, uint32_t ncols, uint64_t *bytes) {
    if (!bytes) return fail(MGV_E_INVALID, "null bytes");
    if (fanout > kMaxFanout) return fail(MGV_E_INVALID, "fan-out above %u", kMaxFanout);
    return plan_layout(n, fanout, ncols, nullptr, 0, nullptr, bytes);
}

int mgv_hashjoin_probe_partitioned(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, int64_t row_base,
                                   uint32_t flags, uint32_t fanout, void *ws, uint64_t ws_bytes, mgv_result *res,
                                   mgv_launch LC, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!slots || !keys || !res || nbuckets == 0) return fail(MGV_E_INVALID, "probe_partitioned: null argument");
    if (n == 0) return MGV_OK;
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "probe_partitioned: n must be < 2^31");
    if (fanout > kMaxFanout) return fail(MGV_E_INVALID, "probe_partitioned: fan-out above %u", kMaxFanout);
    const bool digest = (flags & MGV_PROBE_DIGEST) != 0;
    PartLayout L;
    int rc = plan_layout(n, fanout, digest ? 3 : 2, ws, ws_bytes, &L, nullptr);
    if (rc) return rc;
    hipStream_t s = stream_of(st);
    PartIn in = {pred, bound, keys, vals, row_base, digest ? 1u : 0u, 0u, 0u, res, nullptr};
    if ((rc = partition_rows(L, in, s, scatter_shape(LC)))) return rc;
    if (digest) {
        PProbeSum<false, true, true>::Params p = {0, slots, nbuckets, res};
        return part_probe<PProbeSum<false, true, true>>(p, L, LC, s);
    }
    PProbeSum<false, true, false>::Params p = {0, slots, nbuckets, res};
    if (LC.reserved & MGV_LAUNCH_COOP_PROBE) {
        PProbeSum<false, true, false, OpProbeCoop>::Params pc = {0, slots, nbuckets, res};
        return part_probe<PProbeSum<false, true, false, OpProbeCoop>>(pc, L, LC, s);
    }
    return part_probe<PProbeSum<false, true, false>>(p, L, LC, s);
}

int mgv_hashjoin_probe_groupby_partitioned(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred,
                                           int32_t bound, const int32_t *keys, const int32_t *vals, uint64_t n,
                                           uint64_t *gcnt, int64_t *gsum, uint32_t ngroups, uint32_t fanout,
                                           void *ws, uint64_t ws_bytes, mgv_launch LC, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!slots || !keys || !vals || !gcnt || !gsum || nbuckets == 0 || ngroups == 0)
        return fail(MGV_E_INVALID, "probe_groupby_partitioned: bad argument");
    if (n == 0) return MGV_OK;
    if (n >= (1ull << 31)) return fail(MGV_E_INVALID, "probe_groupby_partitioned: n must be < 2^31");
    if (fanout > kMaxFanout) return fail(MGV_E_INVALID, "probe_groupby_partitioned: fan-out above %u", kMaxFanout);
    PartLayout L;
    int rc = plan_layout(n, fanout, 2, ws, ws_bytes, &L, nullptr);
    if (rc) return rc;
    hipStream_t s = stream_of(st);
    PartIn in = {pred, bound, keys, vals, 0, 0u, 0u, 0u, nullptr, nullptr};
    if ((rc = partition_rows(L, in, s, scatter_shape(LC)))) return rc;
    if (ngroups <= 1024) {
        PProbeGroup<false, true, 1024>::Params p = {0, slots, nbuckets, gcnt, gsum, ngroups};
        return part_probe<PProbeGroup<false, true, 1024>>(p, L, LC, s);
    }
    if (ngroups <= 4096) {
        PProbeGroup<false, true, 4096>::Params p = {0, slots, nbuckets, gcnt, gsum, ngroups};
        return part_probe<PProbeGroup<false, true, 4096>>(p, L, LC, s);
    }
    PProbeGroup<false, true, 0>::Params p = {0, slots, nbuckets, gcnt, gsum, ngroups};
    return part_probe<PProbeGroup<false, true, 0>>(p, L, LC, s);
}

int mgv_debug_xcc_histogram(uint32_t *out8, uint32_t grid, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!out8 || grid == 0) return fail(MGV_E_INVALID, "bad argument");
    hipLaunchKernelGGL(k_xcc_histogram, dim3(grid), dim3(64), 0, stream_of(st), out8);
    return check_launch("k_xcc_histogram");
}

}  // extern "C"
