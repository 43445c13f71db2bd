// mgv_stream.hip -- the device->host completion ring and the streaming
// select into rotating pool buffers.
//
// The reference streams a select's output to the host through gpu_to_cpu:
// each warp appends its compacted rows to a shared output buffer
// (gpu_to_cpu::consume_warp, src/operators/gpu_to_cpu.cu:61-105); a warp
// whose append does not fit swaps in a fresh pool buffer
// (buffer_manager::acquire_buffer_blocked_to, src/buffer_manager.cuh:
// 147-176) and throws the full one onto a pinned ring (throw2cpu, 42-58),
// which a host thread drains (catcher, 149-165).  Here the same flow runs on
// 64-lane waves with the structures of mgv_pool.hpp: one reservation per
// wave on the buffer's count, a CAS of a {generation, index} word to swap buffers, a buffer handed
// to the host only when all its reserved rows are written, a ticketed ring,
// and buffers returned by the host through the pool's return ring.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "mgv_ops.hpp"
#include "mgv_pool.hpp"

namespace mgv {

constexpr int kStreamBlock = 256;
constexpr uint32_t kStage = 512;             // per-wave staging ring (rows)
// An append waits at most this long for a free buffer (REALTIME counter at
// 100 MHz) before its rows are dropped and counted: a host that stopped
// handing buffers back cannot hang the kernel.
constexpr uint64_t kAppendWaitTicks = 10ull * 100000000ull;   // 10 s

struct StreamArgs {
    DevPool *pool;
    StreamCtl *ctl;
    RingDev ring;
    uint32_t cap;           // rows per buffer
};

// Appends rows [head, head + n) (mod kStage) of the wave's staging ring to
// the current output buffer (the whole wave; n <= 256, wave-uniform).
__device__ void stream_append(const StreamArgs &a, const int32_t *S, uint32_t head, uint32_t n) {
    const uint32_t l = lane_id();
    unsigned long long tried = ~0ull;   // the buffer word this wave already failed to reserve in
    uint64_t deadline = 0;              // set at the first wait
    for (;;) {
        unsigned long long w = 0;
        if (l == 0) w = __hip_atomic_load(&a.ctl->cur, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        w = first_lane_u64(w);
        const uint32_t b = (uint32_t)w;
        if (b != kNull && w != tried) {
            DevBuffer *B = &a.pool->bufs[b];
            uint32_t off = 0;
            if (l == 0) off = buffer_reserve(B, w, n, a.cap);
            off = first_lane_u32(off);
            if (off == kBusy) {   // heavy contention on the count: read the word again
                const uint64_t now = __builtin_amdgcn_s_memrealtime();
                if (!deadline) deadline = now + kAppendWaitTicks;
                if (now >= deadline) {
                    if (l == 0) atomicAdd(&a.ctl->dropped, (unsigned long long)n);
                    return;
                }
                continue;
            }
            if (off < kBusy && (uint64_t)off + n <= a.cap) {
                int32_t *dst = B->data + off;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t i = l + (uint32_t)k * kWave;
                    if (i < n) dst[i] = S[(head + i) & (kStage - 1)];
                }
                __threadfence_system();   // the rows (possibly in host memory) before the count
                if (l == 0) buffer_written(a.pool, a.ctl, a.ring, b, n);
                return;
            }
            // the first reservation that did not fit retires the buffer
            if (off <= a.cap && l == 0) buffer_seal(a.pool, a.ctl, a.ring, b, off);
            tried = w;
        }
        unsigned long long r = 0;
        if (l == 0) r = output_swap(a.pool, a.ctl, w);
        r = first_lane_u64(r);
        if ((uint32_t)r == kNull) {   // no free buffer: back off, up to the deadline
            const uint64_t now = __builtin_amdgcn_s_memrealtime();
            if (!deadline) deadline = now + kAppendWaitTicks;
            if (now >= deadline) {
                if (l == 0) atomicAdd(&a.ctl->dropped, (unsigned long long)n);
                return;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) __builtin_amdgcn_s_sleep(127);
        }
    }
}

// Loads rows [i0, i0 + 4) of a column, zeros past n.
__device__ __forceinline__ v4i load_quad_n(const int32_t *p, uint64_t i0, uint64_t n) {
    if (i0 + 3 < n) return load_stream4(p + i0);
    v4i q = {0, 0, 0, 0};
    if (i0 < n) q.x = load_stream1(p + i0);
    if (i0 + 1 < n) q.y = load_stream1(p + i0 + 1);
    if (i0 + 2 < n) q.z = load_stream1(p + i0 + 2);
    return q;
}

// Every wave takes chunks of 256 rows (4 per lane), compacts the survivors
// into its LDS staging ring, and appends 256 rows at a time; the rest is
// appended when the wave has no chunk left.
__global__ __launch_bounds__(kStreamBlock) void k_select_stream(const int32_t *pred, const int32_t *val, uint64_t n,
                                                                int32_t bound, StreamArgs a) {
    __shared__ int32_t stage[kStreamBlock / kWave][kStage];
    int32_t *S = stage[wave_id()];
    const uint32_t l = lane_id();
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;
    const uint64_t waves = (uint64_t)gridDim.x * (kStreamBlock / kWave);
    uint32_t head = 0, fill = 0;   // wave-uniform
    for (uint64_t c = (uint64_t)blockIdx.x * (kStreamBlock / kWave) + wave_id(); c < nchunks; c += waves) {
        const uint64_t i0 = c * kChunk + 4ull * l;
        const v4i p = load_quad_n(pred, i0, n);
        const v4i v = val == pred ? p : load_quad_n(val, i0, n);
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) m |= (uint32_t)(i0 + j < n && q_get(p, j) <= bound) << j;
        // exclusive prefix of the per-lane survivor counts (0..4), bit by bit
        const uint32_t cnt = (uint32_t)__builtin_popcount(m);
        const uint64_t b1 = ballot(cnt & 1u), b2 = ballot(cnt & 2u), b4 = ballot(cnt & 4u);
        uint32_t pos = head + fill + mbcnt(b1) + 2 * mbcnt(b2) + 4 * mbcnt(b4);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if ((m >> j) & 1u) S[(pos++) & (kStage - 1)] = q_get(v, j);
        fill += popc64(b1) + 2 * popc64(b2) + 4 * popc64(b4);
        wave_sync();
        if (fill >= (uint32_t)kChunk) {
            stream_append(a, S, head, kChunk);
            head = (head + kChunk) & (kStage - 1);
            fill -= kChunk;
            wave_sync();   // the staged rows are read before they are overwritten
        }
    }
    if (fill) stream_append(a, S, head, fill);
}

__global__ void k_stream_open(StreamCtl *c) {
    if (threadIdx.x == 0) {
        c->cur = kNull;
        c->rows = 0;
        c->dropped = 0;
        c->buffers = 0;
    }
}

// gpu_to_cpu::at_close: retire the current buffer (every append before this
// kernel is complete, so the seal is the last step and hands it on), then
// publish the end-of-stream entry.
__global__ void k_stream_close(StreamArgs a) {
    if (threadIdx.x != 0) return;
    const unsigned long long w = __hip_atomic_load(&a.ctl->cur, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t b = (uint32_t)w;
    if (b != kNull) {
        // a reservation of cap + 1 rows never fits: if the buffer is not
        // retired yet, it is the first failure and seals it
        uint32_t off = kBusy;
        while (off == kBusy) off = buffer_reserve(&a.pool->bufs[b], w, a.cap + 1, a.cap);
        if (off <= a.cap) buffer_seal(a.pool, a.ctl, a.ring, b, off);
        a.ctl->cur = kNull;
    }
    __threadfence();
    mgv_result r;
    memset(&r, 0, sizeof(r));
    r.count = a.ctl->rows;
    r.overflow = a.ctl->dropped;
    r.reserved[0] = a.ctl->buffers;
    ring_push(a.ring, ~0ull, a.ctl->rows, &r);
}

__global__ void k_ring_emit(RingDev r, uint64_t id, uint64_t value, const mgv_result *res) {
    if (threadIdx.x == 0) ring_push(r, id, value, res);
}

__global__ void k_ring_emit_results(RingDev r, const uint64_t *ids, const mgv_result *res, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) ring_push(r, ids[i], i, res + i);
}

}  // namespace mgv

using namespace mgv;

struct mgv_ring {
    RingSlot *slots;              // pinned coherent host memory
    uint32_t nslots;
    uint64_t front;               // next ticket the host takes
    unsigned long long *d_tail;   // device ticket counter
    unsigned int *lost;           // pinned coherent host memory
    RingDev dev() const { return RingDev{slots, nslots, d_tail, lost}; }
};

struct mgv_stream_out {
    mgv_pool *pool;
    mgv_ring *ring;
    StreamCtl *ctl;               // device memory
    StreamArgs args() const { return StreamArgs{pool->dev, ctl, ring->dev(), pool->host.rows}; }
};

extern "C" {

int mgv_ring_create(uint32_t nslots, mgv_ring **out) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!out || nslots == 0) return fail(MGV_E_INVALID, "mgv_ring_create: bad argument");
    mgv_ring *r = new (std::nothrow) mgv_ring();
    if (!r) return fail(MGV_E_CAPACITY, "out of host memory");
    int rc = hip_status(hipHostMalloc((void **)&r->slots, (size_t)nslots * sizeof(RingSlot), hipHostMallocCoherent),
                        "hipHostMalloc");
    if (!rc) rc = hip_status(hipHostMalloc((void **)&r->lost, 64, hipHostMallocCoherent), "hipHostMalloc");
    if (!rc) rc = hip_status(hipMalloc((void **)&r->d_tail, 64), "hipMalloc");
    if (!rc) rc = hip_status(hipMemset(r->d_tail, 0, 64), "hipMemset");
    if (rc) {
        if (r->slots) (void)hipHostFree(r->slots);
        if (r->lost) (void)hipHostFree(r->lost);
        if (r->d_tail) (void)hipFree(r->d_tail);
        delete r;
        return rc;
    }
    memset(r->slots, 0, (size_t)nslots * sizeof(RingSlot));
    *r->lost = 0;
    r->nslots = nslots;
    r->front = 0;
    *out = r;
    return MGV_OK;
}

int mgv_ring_destroy(mgv_ring *r) {
    if (!r) return MGV_OK;
    (void)hipDeviceSynchronize();
    (void)hipHostFree(r->slots);
    (void)hipHostFree(r->lost);
    (void)hipFree(r->d_tail);
    delete r;
    return MGV_OK;
}

int mgv_ring_emit(mgv_ring *r, uint64_t id, uint64_t value, const mgv_result *res, mgv_stream_t st) {
    if (!r) return fail(MGV_E_INVALID, "mgv_ring_emit: null ring");
    hipLaunchKernelGGL(k_ring_emit, dim3(1), dim3(64), 0, stream_of(st), r->dev(), id, value, res);
    return check_launch("k_ring_emit");
}

int mgv_ring_emit_results(mgv_ring *r, const uint64_t *ids, const mgv_result *res, uint32_t n, mgv_stream_t st) {
    if (!r || (n && (!ids || !res))) return fail(MGV_E_INVALID, "mgv_ring_emit_results: bad argument");
    if (n == 0) return MGV_OK;
    hipLaunchKernelGGL(k_ring_emit_results, dim3(1), dim3(256), 0, stream_of(st), r->dev(), ids, res, n);
    return check_launch("k_ring_emit_results");
}

int mgv_ring_poll(mgv_ring *r, mgv_ring_entry *out) {
    if (!r || !out) return fail(MGV_E_INVALID, "mgv_ring_poll: bad argument");
    if (__atomic_load_n(r->lost, __ATOMIC_ACQUIRE) != 0u)
        return fail(MGV_E_CAPACITY, "mgv_ring_poll: %u entries were lost on a full ring", *r->lost);
    RingSlot *s = &r->slots[r->front % r->nslots];
    if (__atomic_load_n(&s->flag, __ATOMIC_ACQUIRE) == 0u) return 1;
    memcpy(out, (const void *)&s->e, sizeof(*out));
    __atomic_store_n(&s->flag, 0u, __ATOMIC_RELEASE);
    ++r->front;
    return MGV_OK;
}

int mgv_stream_out_create(mgv_pool *pool, mgv_ring *ring, mgv_stream_out **out) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!pool || !ring || !out) return fail(MGV_E_INVALID, "mgv_stream_out_create: null argument");
    if (ring->nslots < pool->host.nbuffers + 2)
        return fail(MGV_E_INVALID, "mgv_stream_out_create: the ring needs %u slots (nbuffers + 2), has %u",
                    pool->host.nbuffers + 2, ring->nslots);
    mgv_stream_out *o = new (std::nothrow) mgv_stream_out();
    if (!o) return fail(MGV_E_CAPACITY, "out of host memory");
    o->pool = pool;
    o->ring = ring;
    int rc = hip_status(hipMalloc((void **)&o->ctl, sizeof(StreamCtl)), "hipMalloc");
    if (!rc) {
        StreamCtl c = {(unsigned long long)kNull, 0, 0, 0};
        rc = hip_status(hipMemcpy(o->ctl, &c, sizeof(c), hipMemcpyHostToDevice), "hipMemcpy");
    }
    if (rc) {
        (void)hipFree(o->ctl);
        delete o;
        return rc;
    }
    *out = o;
    return MGV_OK;
}

int mgv_stream_out_destroy(mgv_stream_out *o) {
    if (!o) return MGV_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(o->ctl);
    delete o;
    return MGV_OK;
}

int mgv_stream_out_open(mgv_stream_out *o, mgv_stream_t st) {
    if (!o) return fail(MGV_E_INVALID, "mgv_stream_out_open: null argument");
    hipLaunchKernelGGL(k_stream_open, dim3(1), dim3(64), 0, stream_of(st), o->ctl);
    return check_launch("k_stream_open");
}

int mgv_stream_out_close(mgv_stream_out *o, mgv_stream_t st) {
    if (!o) return fail(MGV_E_INVALID, "mgv_stream_out_close: null argument");
    hipLaunchKernelGGL(k_stream_close, dim3(1), dim3(64), 0, stream_of(st), o->args());
    return check_launch("k_stream_close");
}

int mgv_select_stream(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_stream_out *o,
                      mgv_launch LC, mgv_stream_t st) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!o || !pred || !val) return fail(MGV_E_INVALID, "mgv_select_stream: null argument");
    if (n == 0) return MGV_OK;
    if (!aligned16(pred) || !aligned16(val)) return fail(MGV_E_ALIGN, "mgv_select_stream: columns not 16-byte aligned");
    const uint64_t chunks = (n + kChunk - 1) / kChunk;
    uint64_t grid = LC.grid ? LC.grid : (chunks + kStreamBlock / kWave - 1) / (kStreamBlock / kWave);
    const uint64_t cap = (uint64_t)current_state()->cus * 8;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(k_select_stream, dim3((uint32_t)grid), dim3(kStreamBlock), 0, stream_of(st), pred, val, n,
                       bound, o->args());
    return check_launch("k_select_stream");
}

}  // extern "C"
