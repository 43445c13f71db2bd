// mgv_pool.hip -- device buffer pool, lock-free device stack and the
// device->host result channel.
//
// Buffer pool: nbuffers device buffers of `rows` int32 each (one arena), a
// per-buffer descriptor {cnt, data}, and a lock-free stack of free buffer
// indices with the reference pool's counters (producers, available).
// Replaces buffer_pool<T> (src/buffer_pool.cuh:35-170), the device part of
// buffer_manager<int32_t> (src/buffer_manager.cuh:40-181) and its stacks
// (src/lockfree_stack.cuh:19-208, src/threadsafe_device_stack.cuh:19-87).
//
// The stack is a Treiber stack over buffer indices whose head packs {index,
// ABA tag} into one 64-bit word updated with a 64-bit CAS.  Stack operations
// are done by one lane (the caller elects it), like the reference's, and no
// wave ever waits on a lock held by another lane: a failed CAS just retries
// with the value it returned, so lanes of one wave cannot deadlock.
//
// The pool's structures and device operations are in mgv_pool.hpp; buffers
// come back from the host either with a kernel (mgv_pool_release) or,
// without one, through the pool's return ring (mgv_pool_release_host).
//
// Result channel: a pinned host ring of slots {flag, mgv_result}.  A kernel
// copies a device result into a slot and publishes it with a system-scope
// release; a host thread polls the flag.  This is the reduced form of
// gpu_to_cpu's pinned ring (src/operators/gpu_to_cpu.cu:42-58,149-165).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "mgv_pool.hpp"

namespace mgv {

__global__ void k_pool_init(DevPool *p) {
    // one thread: link all buffers into the stack
    for (uint32_t i = 0; i < p->nbuffers; ++i) {
        p->next[i] = (i + 1 < p->nbuffers) ? i + 1 : kNull;
        p->bufs[i].cnt = 0;
        p->bufs[i].gen = 0;
        p->bufs[i].done = 0;
        p->bufs[i].final_rows = 0;
        p->owner[i] = 0;
    }
    p->head = p->nbuffers ? 0ull : (unsigned long long)kNull;
    p->producers = 0;
    p->available = p->nbuffers;
    p->ret_tail = 0;
}

__global__ void k_pool_acquire(DevPool *p, int32_t *out) {
    if (threadIdx.x == 0) {
        const int32_t b = pool_acquire(p);
        if (b >= 0) p->bufs[b].cnt = 0;   // an acquired buffer starts empty
        *out = b;
    }
}

__global__ void k_pool_release(DevPool *p, uint32_t i) {
    if (threadIdx.x == 0) pool_release(p, i);
}

__global__ void k_pool_producer(DevPool *p, int32_t delta) {
    if (threadIdx.x == 0) {
        if (delta > 0) atomicAdd(&p->producers, 1u);
        else atomicSub(&p->producers, 1u);
    }
}

__global__ void k_pool_stats(DevPool *p, uint32_t *out) {
    if (threadIdx.x == 0) {
        out[0] = p->available;
        out[1] = p->producers;
        uint32_t n = 0;
        for (uint32_t i = (uint32_t)p->head; i != kNull && n <= p->nbuffers; i = p->next[i]) ++n;
        out[2] = n;   // buffers on the stack (quiescent pool)
        out[3] = (uint32_t)p->ret_tail;   // return-ring entries taken (low 32 bits)
    }
}

// Stress: every wave repeatedly acquires a buffer with one elected lane,
// checks exclusive ownership with an atomic flag, writes a stamp, and
// releases it.  ops[0] = successful acquire/release pairs, ops[1] = ownership
// violations, ops[2] = RETRY results, ops[3] = END results.  Every loop is
// bounded by `iters`, so every wave exits.
__global__ void k_pool_stress(DevPool *p, uint32_t iters, unsigned long long *ops) {
    const bool leader = lane_id() == 0;
    unsigned long long ok = 0, bad = 0, retry = 0, end = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        int32_t b = 0;
        if (leader) b = pool_acquire(p);
        b = __builtin_amdgcn_readfirstlane(b);
        if (b == MGV_POOL_RETRY) { ++retry; continue; }
        if (b == MGV_POOL_END) { ++end; continue; }
        if (leader) {
            if (atomicExch(&p->owner[b], 1u) != 0u) ++bad;
            p->bufs[b].data[0] = (int32_t)(blockIdx.x * 64 + threadIdx.x / 64);
            if (atomicExch(&p->owner[b], 0u) != 1u) ++bad;
            pool_release(p, (uint32_t)b);
        }
        ++ok;
    }
    if (leader) {
        atomicAdd(&ops[0], ok);
        atomicAdd(&ops[1], bad);
        atomicAdd(&ops[2], retry);
        atomicAdd(&ops[3], end);
    }
}

// Select materialisation into one pool buffer: buffer::try_write /
// try_partial_final_write (src/buffer.cuh:92-134): every wave reserves its
// survivors' slots with one atomicAdd on the buffer count; rows beyond the
// capacity are not written (the count keeps growing so the caller sees how
// many were produced).
__global__ void k_select_to_buffer(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, DevPool *p,
                                   uint32_t bi) {
    DevBuffer *b = &p->bufs[bi];
    const uint32_t cap = p->rows;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x - lane_id()) ; i0 < n;
         i0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = i0 + lane_id();
        const bool pass = i < n && pred[i] <= bound;
        const uint64_t bal = ballot(pass);
        if (!bal) continue;
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(&b->cnt, popc64(bal));
        base = __builtin_amdgcn_readfirstlane(base);
        const uint32_t pos = base + mbcnt(bal);
        if (pass && pos < cap) b->data[pos] = val[i];
    }
}

// gpu_to_cpu, reduced: copy a device result into a pinned host slot and
// publish it with a system-scope release on the slot flag.
struct HostSlot {
    uint32_t flag;
    uint32_t seq;
    uint64_t pad;
    mgv_result res;
};

__global__ void k_emit_result(const mgv_result *src, HostSlot *slot, uint32_t seq) {
    if (threadIdx.x < sizeof(mgv_result) / 8)
        reinterpret_cast<volatile uint64_t *>(&slot->res)[threadIdx.x] = reinterpret_cast<const uint64_t *>(src)[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(&slot->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&slot->flag, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace mgv

using namespace mgv;

struct mgv_channel {
    HostSlot *slots;       // pinned host ring
    uint32_t nslots;
};

extern "C" {

static void pool_free(mgv_pool *P) {
    (void)hipFree(P->dev);
    if (P->host_arena) {
        if (P->arena) (void)hipHostFree(P->arena);
    } else {
        (void)hipFree(P->arena);
    }
    (void)hipFree(P->host.next);
    (void)hipFree(P->host.bufs);
    (void)hipFree(P->host.owner);
    if (P->acq_slot) (void)hipHostFree(P->acq_slot);
    if (P->stats) (void)hipHostFree(P->stats);
    if (P->ret) (void)hipHostFree(P->ret);
    delete P->ret_mu;
    delete P;
}

// A pool whose buffers are device memory, or pinned coherent host memory
// (host_arena: kernels write the rows over PCIe, the host reads them in
// place).
static int pool_create(uint32_t nbuffers, uint32_t rows, bool host_arena, mgv_pool **out) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!out || nbuffers == 0 || nbuffers >= kNull - 1 || rows == 0 || rows % 256 || rows >= kSeal)
        return fail(MGV_E_INVALID, "mgv_pool_create: need nbuffers > 0 and rows a positive multiple of 256 below 2^31");
    mgv_pool *P = new (std::nothrow) mgv_pool();
    if (!P) return fail(MGV_E_CAPACITY, "out of host memory");
    P->ret_mu = new (std::nothrow) std::mutex();
    int rc = P->ret_mu ? 0 : fail(MGV_E_CAPACITY, "out of host memory");
    const uint32_t nret = nbuffers + 1;   // every buffer fits at once: the host never finds it full
    do {
        if (rc) break;
        if ((rc = hip_status(hipMalloc(&P->dev, sizeof(DevPool)), "hipMalloc"))) break;
        P->host_arena = host_arena;
        if (host_arena) {
            if ((rc = hip_status(hipHostMalloc((void **)&P->arena, (size_t)nbuffers * rows * 4, hipHostMallocCoherent),
                                 "hipHostMalloc")))
                break;
        } else if ((rc = hip_status(hipMalloc(&P->arena, (size_t)nbuffers * rows * 4), "hipMalloc"))) {
            break;
        }
        if ((rc = hip_status(hipMalloc(&P->host.next, (size_t)nbuffers * 4), "hipMalloc"))) break;
        if ((rc = hip_status(hipMalloc(&P->host.bufs, (size_t)nbuffers * sizeof(DevBuffer)), "hipMalloc"))) break;
        if ((rc = hip_status(hipMalloc(&P->host.owner, (size_t)nbuffers * 4), "hipMalloc"))) break;
        if ((rc = hip_status(hipHostMalloc((void **)&P->acq_slot, 64, hipHostMallocDefault), "hipHostMalloc"))) break;
        if ((rc = hip_status(hipHostMalloc((void **)&P->stats, 64, hipHostMallocDefault), "hipHostMalloc"))) break;
        if ((rc = hip_status(hipHostMalloc((void **)&P->ret, (size_t)nret * sizeof(ReturnSlot), hipHostMallocCoherent),
                             "hipHostMalloc")))
            break;
        memset(P->ret, 0, (size_t)nret * sizeof(ReturnSlot));
        P->ret_head = 0;
        P->host.nbuffers = nbuffers;
        P->host.rows = rows;
        P->host.head = 0;
        P->host.producers = 0;
        P->host.available = nbuffers;
        P->host.ret = P->ret;
        P->host.nret = nret;
        P->host.ret_tail = 0;
        // descriptors: data pointers into the arena
        DevBuffer *d = new (std::nothrow) DevBuffer[nbuffers];
        if (!d) { rc = fail(MGV_E_CAPACITY, "out of host memory"); break; }
        for (uint32_t i = 0; i < nbuffers; ++i) d[i] = DevBuffer{0, 0, P->arena + (size_t)i * rows, 0, 0};
        rc = hip_status(hipMemcpy(P->host.bufs, d, (size_t)nbuffers * sizeof(DevBuffer), hipMemcpyHostToDevice),
                        "hipMemcpy");
        delete[] d;
        if (rc) break;
        if ((rc = hip_status(hipMemcpy(P->dev, &P->host, sizeof(DevPool), hipMemcpyHostToDevice), "hipMemcpy"))) break;
        hipLaunchKernelGGL(k_pool_init, dim3(1), dim3(1), 0, stream_of(nullptr), P->dev);
        if ((rc = check_launch("k_pool_init"))) break;
        if ((rc = hip_status(hipStreamSynchronize(stream_of(nullptr)), "hipStreamSynchronize"))) break;
        *out = P;
        return MGV_OK;
    } while (0);
    pool_free(P);
    return rc;
}

int mgv_pool_create(uint32_t nbuffers, uint32_t rows, mgv_pool **out) {
    return pool_create(nbuffers, rows, false, out);
}

int mgv_pool_create_host(uint32_t nbuffers, uint32_t rows, mgv_pool **out) {
    return pool_create(nbuffers, rows, true, out);
}

int mgv_pool_destroy(mgv_pool *P) {
    if (!P) return MGV_OK;
    (void)hipDeviceSynchronize();
    pool_free(P);
    return MGV_OK;
}

int mgv_pool_release_host(mgv_pool *P, int32_t buffer) {
    if (!P || buffer < 0 || (uint32_t)buffer >= P->host.nbuffers)
        return fail(MGV_E_INVALID, "mgv_pool_release_host: bad buffer %d", buffer);
    std::lock_guard<std::mutex> lk(*P->ret_mu);
    ReturnSlot *s = &P->ret[P->ret_head % P->host.nret];
    // the ring has a slot per buffer (plus one), so a slot the device has not
    // emptied means a buffer was released twice
    if (__atomic_load_n(&s->flag, __ATOMIC_ACQUIRE) != 0u)
        return fail(MGV_E_CAPACITY, "mgv_pool_release_host: return ring full (buffer %d released twice?)", buffer);
    __atomic_store_n(&s->index, buffer, __ATOMIC_RELAXED);
    __atomic_store_n(&s->flag, 1u, __ATOMIC_RELEASE);
    ++P->ret_head;
    return MGV_OK;
}

int mgv_pool_acquire(mgv_pool *P, int32_t *buffer, mgv_stream_t st) {
    if (!P || !buffer) return fail(MGV_E_INVALID, "mgv_pool_acquire: null argument");
    hipStream_t s = stream_of(st);
    hipLaunchKernelGGL(k_pool_acquire, dim3(1), dim3(64), 0, s, P->dev, P->acq_slot);
    int rc = check_launch("k_pool_acquire");
    if (rc) return rc;
    if ((rc = hip_status(hipStreamSynchronize(s), "hipStreamSynchronize"))) return rc;
    *buffer = *(volatile int32_t *)P->acq_slot;
    return MGV_OK;
}

int mgv_pool_release(mgv_pool *P, int32_t buffer, mgv_stream_t st) {
    if (!P || buffer < 0 || (uint32_t)buffer >= P->host.nbuffers)
        return fail(MGV_E_INVALID, "mgv_pool_release: bad buffer %d", buffer);
    hipLaunchKernelGGL(k_pool_release, dim3(1), dim3(64), 0, stream_of(st), P->dev, (uint32_t)buffer);
    return check_launch("k_pool_release");
}

int mgv_pool_register_producer(mgv_pool *P, mgv_stream_t st) {
    if (!P) return fail(MGV_E_INVALID, "null pool");
    hipLaunchKernelGGL(k_pool_producer, dim3(1), dim3(64), 0, stream_of(st), P->dev, 1);
    return check_launch("k_pool_producer");
}

int mgv_pool_unregister_producer(mgv_pool *P, mgv_stream_t st) {
    if (!P) return fail(MGV_E_INVALID, "null pool");
    hipLaunchKernelGGL(k_pool_producer, dim3(1), dim3(64), 0, stream_of(st), P->dev, -1);
    return check_launch("k_pool_producer");
}

int mgv_pool_buffer(mgv_pool *P, int32_t buffer, int32_t **data, uint32_t **cnt) {
    if (!P || buffer < 0 || (uint32_t)buffer >= P->host.nbuffers)
        return fail(MGV_E_INVALID, "mgv_pool_buffer: bad buffer %d", buffer);
    if (data) *data = P->arena + (size_t)buffer * P->host.rows;
    if (cnt) *cnt = &P->host.bufs[buffer].cnt;
    return MGV_OK;
}

int mgv_pool_info(mgv_pool *P, uint32_t *nbuffers, uint32_t *rows) {
    if (!P) return fail(MGV_E_INVALID, "null pool");
    if (nbuffers) *nbuffers = P->host.nbuffers;
    if (rows) *rows = P->host.rows;
    return MGV_OK;
}

int mgv_pool_stats(mgv_pool *P, uint32_t *available, uint32_t *producers, uint32_t *stacked, mgv_stream_t st) {
    if (!P) return fail(MGV_E_INVALID, "null pool");
    hipStream_t s = stream_of(st);
    hipLaunchKernelGGL(k_pool_stats, dim3(1), dim3(64), 0, s, P->dev, P->stats);
    int rc = check_launch("k_pool_stats");
    if (rc) return rc;
    if ((rc = hip_status(hipStreamSynchronize(s), "hipStreamSynchronize"))) return rc;
    if (available) *available = P->stats[0];
    if (producers) *producers = P->stats[1];
    if (stacked) {
        // buffers on the stack plus those handed back through the return ring
        // and not yet taken by the device
        std::lock_guard<std::mutex> lk(*P->ret_mu);
        *stacked = P->stats[2] + (uint32_t)((uint32_t)P->ret_head - P->stats[3]);
    }
    return MGV_OK;
}

int mgv_pool_stress(mgv_pool *P, uint32_t blocks, uint32_t iters, uint64_t *ops4_dev, mgv_stream_t st) {
    if (!P || !ops4_dev || blocks == 0) return fail(MGV_E_INVALID, "mgv_pool_stress: bad argument");
    hipLaunchKernelGGL(k_pool_stress, dim3(blocks), dim3(256), 0, stream_of(st), P->dev, iters,
                       (unsigned long long *)ops4_dev);
    return check_launch("k_pool_stress");
}

int mgv_select_to_buffer(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_pool *P,
                         int32_t buffer, mgv_stream_t st) {
    if (!P || !pred || !val || buffer < 0 || (uint32_t)buffer >= P->host.nbuffers)
        return fail(MGV_E_INVALID, "mgv_select_to_buffer: bad argument");
    if (n == 0) return MGV_OK;
    uint64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_select_to_buffer, dim3((uint32_t)g), dim3(256), 0, stream_of(st), pred, val, n, bound,
                       P->dev, (uint32_t)buffer);
    return check_launch("k_select_to_buffer");
}

int mgv_channel_create(uint32_t nslots, mgv_channel **out) {
    if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called");
    if (!out || nslots == 0) return fail(MGV_E_INVALID, "mgv_channel_create: bad argument");
    mgv_channel *c = new (std::nothrow) mgv_channel();
    if (!c) return fail(MGV_E_CAPACITY, "out of host memory");
    int rc = hip_status(hipHostMalloc((void **)&c->slots, (size_t)nslots * sizeof(HostSlot), hipHostMallocCoherent),
                        "hipHostMalloc");
    if (rc) {
        delete c;
        return rc;
    }
    memset(c->slots, 0, (size_t)nslots * sizeof(HostSlot));
    c->nslots = nslots;
    *out = c;
    return MGV_OK;
}

int mgv_channel_destroy(mgv_channel *c) {
    if (!c) return MGV_OK;
    (void)hipHostFree(c->slots);
    delete c;
    return MGV_OK;
}

int mgv_channel_emit(mgv_channel *c, uint32_t slot, const mgv_result *res, uint32_t seq, mgv_stream_t st) {
    if (!c || slot >= c->nslots || !res) return fail(MGV_E_INVALID, "mgv_channel_emit: bad argument");
    hipLaunchKernelGGL(k_emit_result, dim3(1), dim3(64), 0, stream_of(st), res, &c->slots[slot], seq);
    return check_launch("k_emit_result");
}

int mgv_channel_poll(mgv_channel *c, uint32_t slot, mgv_result *out, uint32_t *seq) {
    if (!c || slot >= c->nslots) return fail(MGV_E_INVALID, "mgv_channel_poll: bad argument");
    HostSlot *h = &c->slots[slot];
    if (__atomic_load_n(&h->flag, __ATOMIC_ACQUIRE) == 0) return 1;
    if (out) memcpy(out, (const void *)&h->res, sizeof(mgv_result));
    if (seq) *seq = __atomic_load_n(&h->seq, __ATOMIC_RELAXED);
    __atomic_store_n(&h->flag, 0u, __ATOMIC_RELEASE);
    return 0;
}

}  // extern "C"
