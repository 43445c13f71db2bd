// mgv_pool.hpp -- device side of the buffer pool, the device->host ring and
// the rotating output buffer, shared by mgv_pool.hip and mgv_stream.hip.
//
// Pool: a Treiber stack of free buffer indices (64-bit head = {ABA tag,
// index}, one-lane operations) plus a host->device return ring in pinned
// memory through which host threads hand buffers back without a kernel
// launch.  Replaces buffer_pool<T> (src/buffer_pool.cuh:35-170) and the
// device stacks (src/lockfree_stack.cuh:19-208).
//
// Ring: slots in pinned coherent host memory, claimed by a ticket counter in
// device memory and published with a system-scope release; a host thread
// polls them in ticket order.  Replaces gpu_to_cpu's lock-protected ring
// (src/operators/gpu_to_cpu.cu:42-58,149-165): a ticket instead of a lock,
// so no lane ever waits for another lane.
//
// Rotating output: the reference's shared output_buffer with
// buffer_manager::acquire_buffer_blocked_to (src/buffer_manager.cuh:147-176)
// and buffer::try_write (src/buffer.cuh:92-134).  The current buffer is a
// {generation, index} word swapped with a 64-bit CAS (no ABA), and a retired
// buffer reaches the host only once every wave that reserved rows in it has
// written them (the reference hands it over while writers may still be
// writing, buffer.cuh:102).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "mgv_internal.hpp"
#include "mgv_wave.hpp"

namespace mgv {

constexpr uint32_t kNull = 0xffffffffu;
// Added to a buffer's `done` count when it is retired; with capacities below
// 2^31 the count reaches exactly kSeal when the last reserved row is written.
constexpr uint32_t kSeal = 1u << 31;

struct DevBuffer {           // the device-visible part of buffer<int32_t, cap>
    uint32_t cnt;            // rows reserved (failed reservations are not undone)
    uint32_t gen;            // generation: raised each time the buffer is about to become current
    int32_t *data;
    uint32_t done;           // rows written, plus kSeal - final once retired
    uint32_t final_rows;     // rows of a retired buffer
};

struct ReturnSlot {          // pinned host memory: one buffer handed back by the host
    uint32_t flag;           // 1 = holds a buffer index
    int32_t index;
};

struct DevPool {
    unsigned long long head;  // (tag << 32) | index, index kNull = empty
    uint32_t producers;
    uint32_t available;
    uint32_t nbuffers;
    uint32_t rows;
    uint32_t *next;           // [nbuffers]
    DevBuffer *bufs;          // [nbuffers]
    uint32_t *owner;          // [nbuffers] stress-test ownership flags
    ReturnSlot *ret;          // [nret] host->device return ring (pinned)
    uint32_t nret;
    uint32_t pad;
    unsigned long long ret_tail;   // next return ticket the device takes
};

__device__ __forceinline__ unsigned long long ld_head(DevPool *p) {
    return __hip_atomic_load(&p->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// push index i (one lane)
__device__ __forceinline__ void stack_push(DevPool *p, uint32_t i) {
    unsigned long long old = ld_head(p);
    for (;;) {
        __hip_atomic_store(&p->next[i], (uint32_t)old, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nw = (((old >> 32) + 1) << 32) | i;
        // release: next[i] is visible before the new head
        if (__hip_atomic_compare_exchange_strong(&p->head, &old, nw, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// pop (one lane): index or kNull when empty
__device__ __forceinline__ uint32_t stack_try_pop(DevPool *p) {
    unsigned long long old = __hip_atomic_load(&p->head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint32_t i = (uint32_t)old;
        if (i == kNull) return kNull;
        const uint32_t n = __hip_atomic_load(&p->next[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nw = (((old >> 32) + 1) << 32) | n;
        if (__hip_atomic_compare_exchange_strong(&p->head, &old, nw, __ATOMIC_ACQUIRE, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return i;
    }
}

// Takes the oldest buffer the host handed back through the return ring (one
// lane): its index, or kNull when the ring holds none.  The ticket is taken
// with a CAS on the monotonic tail, so a stale view of the tail never takes
// a slot twice; the slot is cleared after its index was read.
__device__ __forceinline__ uint32_t take_returned(DevPool *p) {
    if (!p->ret || !p->nret) return kNull;
    unsigned long long t = __hip_atomic_load(&p->ret_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = 0; k < 8; ++k) {
        ReturnSlot *s = &p->ret[t % p->nret];
        if (__hip_atomic_load(&s->flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) == 0u) return kNull;
        if (__hip_atomic_compare_exchange_strong(&p->ret_tail, &t, t + 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            const int32_t i = __hip_atomic_load(&s->index, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&s->flag, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            return (uint32_t)i < p->nbuffers ? (uint32_t)i : kNull;
        }
        // t now holds the current tail: look at that slot
    }
    return kNull;
}

// acquire_buffer_blocked_unsafe (src/buffer_pool.cuh:83-96): at most 100
// attempts; a buffer index, MGV_POOL_END when no producer is registered and
// no buffer is available, or MGV_POOL_RETRY when the attempts ran out.
// Buffers the host handed back are taken when the stack is empty.  The
// buffer's counts are left as they are: the caller resets what it uses.
__device__ __forceinline__ int32_t pool_acquire(DevPool *p) {
    for (int tries = 0; tries < 100; ++tries) {
        uint32_t i = stack_try_pop(p);
        if (i != kNull) atomicSub(&p->available, 1u);
        else i = take_returned(p);
        if (i != kNull) return (int32_t)i;
        const uint32_t prod = __hip_atomic_load(&p->producers, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t avail = __hip_atomic_load(&p->available, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the return ring was just found empty as well
        if (prod == 0 && avail == 0) return MGV_POOL_END;
    }
    return MGV_POOL_RETRY;
}

// release_buffer (src/buffer_pool.cuh:125-130)
__device__ __forceinline__ void pool_release(DevPool *p, uint32_t i) {
    atomicAdd(&p->available, 1u);
    stack_push(p, i);
}

// ------------------------------------------------------------------ ring

struct RingSlot {            // pinned coherent host memory, 96 bytes
    uint32_t flag;           // 1 = holds an entry the host has not taken
    uint32_t pad;
    mgv_ring_entry e;
    uint64_t pad2;
};
static_assert(sizeof(mgv_ring_entry) == 80, "ring entry layout");
static_assert(sizeof(RingSlot) == 96, "ring slot layout");

struct RingDev {
    RingSlot *slots;                 // [nslots], pinned host memory
    uint32_t nslots;
    unsigned long long *tail;        // ticket counter (device memory)
    unsigned int *lost;              // entries dropped on a full ring (pinned host memory)
};

// Publishes one entry (one lane).  The ticket fixes the slot; a slot the
// host has not emptied yet means the ring is full: the entry is counted as
// lost (the host's poll reports it) instead of waiting.
__device__ __forceinline__ bool ring_push(const RingDev &r, uint64_t id, uint64_t value, const mgv_result *res) {
    const unsigned long long t = atomicAdd(r.tail, 1ull);
    RingSlot *s = &r.slots[t % r.nslots];
    if (__hip_atomic_load(&s->flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
        __hip_atomic_fetch_add(r.lost, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return false;
    }
    volatile uint64_t *dst = reinterpret_cast<volatile uint64_t *>(&s->e);
    dst[0] = id;
    dst[1] = value;
    const uint64_t *src = reinterpret_cast<const uint64_t *>(res);
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[2 + k] = res ? src[k] : 0ull;
    __threadfence_system();
    __hip_atomic_store(&s->flag, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    return true;
}

// ------------------------------------------------------------------ rotating output buffer

struct StreamCtl {                  // device memory
    unsigned long long cur;         // (gen << 32) | index of the current output buffer; kNull: none yet
    unsigned long long rows;        // rows handed to the host
    unsigned long long dropped;     // rows no buffer could take in time
    unsigned long long buffers;     // buffers handed to the host
};

// A buffer is retired: every row reserved in it is written.  Its rows go to
// the host through the ring; an empty buffer goes straight back to the pool.
__device__ __forceinline__ void buffer_retire(DevPool *p, StreamCtl *so, const RingDev &r, uint32_t b, uint32_t rows) {
    if (rows == 0) {
        pool_release(p, b);
        return;
    }
    if (ring_push(r, b, rows, nullptr)) {
        atomicAdd(&so->rows, (unsigned long long)rows);
        atomicAdd(&so->buffers, 1ull);
    } else {
        atomicAdd(&so->dropped, (unsigned long long)rows);
        pool_release(p, b);
    }
}

// The first reservation that did not fit (at offset v <= capacity) retires
// the buffer: every earlier reservation fit, every later one fails, so the
// buffer holds exactly v rows.  Whichever of the sealer and the writers
// brings `done` to kSeal is the last and hands the buffer on.
__device__ __forceinline__ void buffer_seal(DevPool *p, StreamCtl *so, const RingDev &r, uint32_t b, uint32_t v) {
    __hip_atomic_store(&p->bufs[b].final_rows, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t add = kSeal - v;
    const uint32_t old = __hip_atomic_fetch_add(&p->bufs[b].done, add, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old + add == kSeal) buffer_retire(p, so, r, b, v);
}

// A writer's n rows are written (the caller fenced them).
__device__ __forceinline__ void buffer_written(DevPool *p, StreamCtl *so, const RingDev &r, uint32_t b, uint32_t n) {
    const uint32_t old = __hip_atomic_fetch_add(&p->bufs[b].done, n, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old + n == kSeal)
        buffer_retire(p, so, r, b,
                      __hip_atomic_load(&p->bufs[b].final_rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// buffer::try_write's reservation (src/buffer.cuh:92-105), one lane: n rows
// in buffer b, which the caller saw current under output word w.  The
// buffer's {cnt, gen} pair is one 64-bit word, so a reservation only lands
// in the incarnation of the buffer that w named: a caller holding a stale
// word (the buffer was retired, handed back and made current again under a
// newer tag) gets kStale and never touches the new incarnation's count.
// Returns the offset (a reservation that does not fit still counts: the
// first such one, at an offset <= cap, must seal the buffer), kNull when the
// buffer is already past its capacity, kStale, or kBusy when the CAS kept
// failing.
constexpr uint32_t kStale = 0xfffffffeu;
constexpr uint32_t kBusy = 0xfffffffdu;

__device__ __forceinline__ uint32_t buffer_reserve(DevBuffer *B, unsigned long long w, uint32_t n, uint32_t cap) {
    unsigned long long *word = reinterpret_cast<unsigned long long *>(&B->cnt);
    unsigned long long s = __hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = 0; k < 256; ++k) {
        if ((uint32_t)(s >> 32) != (uint32_t)(w >> 32)) return kStale;
        const uint32_t c = (uint32_t)s;
        if (c > cap) return kNull;
        if (__hip_atomic_compare_exchange_strong(word, &s, s + n, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return c;
    }
    return kBusy;
}

// acquire_buffer_blocked_to (src/buffer_manager.cuh:147-176), one lane:
// replaces the output buffer the caller saw (`seen`) by a fresh pool buffer
// with a CAS of the output word {gen, index}.  The acquirer owns the fresh
// buffer until the CAS, so it gives it a new generation (its own count plus
// one) and an empty count first; the word names that one incarnation only,
// so the output word never repeats and a stale word never matches a later
// incarnation.  If another wave swapped first the buffer goes back to the
// pool (no word ever named it under that generation, so nothing can reserve
// in it).  Returns the current word (the caller's new buffer, or the one
// another wave installed first), or kNull in the low half when no buffer was
// free within the pool's attempts (the caller retries later; nothing here
// waits on the host).
__device__ __forceinline__ unsigned long long output_swap(DevPool *p, StreamCtl *so, unsigned long long seen) {
    unsigned long long now = __hip_atomic_load(&so->cur, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (now != seen) return now;
    const int32_t nb = pool_acquire(p);
    if (nb < 0) return (seen & ~0xffffffffull) | kNull;
    DevBuffer *B = &p->bufs[nb];
    const unsigned long long gen =
        (unsigned long long)(__hip_atomic_load(&B->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u);
    __hip_atomic_store(&B->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&B->final_rows, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(&B->cnt), gen << 32, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long nw = (gen << 32) | (uint32_t)nb;
    if (__hip_atomic_compare_exchange_strong(&so->cur, &seen, nw, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                             __HIP_MEMORY_SCOPE_AGENT))
        return nw;
    pool_release(p, (uint32_t)nb);   // another wave swapped first: seen now holds its word
    return seen;
}

}  // namespace mgv

// Host handle of a pool.
struct mgv_pool {
    mgv::DevPool *dev;          // device struct
    mgv::DevPool host;          // host copy of the pointers
    int32_t *arena;
    bool host_arena;            // buffers in pinned host memory
    int32_t *acq_slot;          // pinned host int
    uint32_t *stats;            // pinned host uint[4]
    mgv::ReturnSlot *ret;       // pinned coherent return ring [host.nret]
    uint64_t ret_head;          // next return ticket the host writes (under ret_mu)
    std::mutex *ret_mu;
};
