// mgv_wave.hpp -- wave64 primitives for gfx950 (CDNA4).
//
// The reference hard-codes 32-lane warps (common.cuh:25) and reads the lane id
// through inline PTX (common.cuh:211-215); here every helper is written for
// 64-lane wavefronts: 64-bit ballots, mbcnt prefix counts, DPP reductions and
// readlane/readfirstlane broadcasts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgv {

constexpr int kWave = 64;            // lanes per wavefront on CDNA
constexpr int kQuad = 4;             // rows per lane per chunk
constexpr int kChunk = kWave * kQuad;  // rows per wave chunk (4 * WARPSIZE in the reference)

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / kWave; }

// XCC (XCD) id of the executing wave: HW_REG_XCC_ID (hwreg 20), bits [2:0]
// (8 XCDs).  Used only to choose which work a wave takes first: results never
// depend on it.
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Number of set bits of `mask` in lanes strictly below this lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

__device__ __forceinline__ uint32_t first_lane_u32(uint32_t x) {
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Compiler-level ordering point for wave-synchronous LDS exchanges: LDS
// instructions of one wave are processed in order, this keeps the compiler
// from reordering them across the exchange.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// DPP controls (GFX9 encoding).
constexpr int kDppQuadXor1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppRowMirror = 0x140;  // row_mirror

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    uint32_t lo = dpp_u32<CTRL>((uint32_t)v);
    uint32_t hi = dpp_u32<CTRL>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Wave-wide sum (two's complement, mod 2^64).  All 64 lanes must be active.
// Four DPP steps leave every lane of a 16-lane row holding the row sum; the
// four row sums are combined with readlane.  The result is wave-uniform.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    v += dpp_u64<kDppQuadXor1>(v);
    v += dpp_u64<kDppQuadXor2>(v);
    v += dpp_u64<kDppHalfMirror>(v);
    v += dpp_u64<kDppRowMirror>(v);
    uint64_t r = 0;
#pragma unroll
    for (int row = 0; row < 4; ++row) {
        uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, row * 16);
        uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), row * 16);
        r += ((uint64_t)hi << 32) | lo;
    }
    return r;
}

__device__ __forceinline__ uint64_t wave_xor_u64(uint64_t v) {
    v ^= dpp_u64<kDppQuadXor1>(v);
    v ^= dpp_u64<kDppQuadXor2>(v);
    v ^= dpp_u64<kDppHalfMirror>(v);
    v ^= dpp_u64<kDppRowMirror>(v);
    uint64_t r = 0;
#pragma unroll
    for (int row = 0; row < 4; ++row) {
        uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, row * 16);
        uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), row * 16);
        r ^= ((uint64_t)hi << 32) | lo;
    }
    return r;
}

// Streaming loads that should not displace hash tables from L2 / MALL.  The
// pointers are cast to the global address space: columns are device or pinned
// host memory, never LDS or scratch, and a pointer the compiler cannot place
// (one read from a query descriptor) would otherwise become a flat load,
// which also counts against the LDS counter, so every LDS wait would drain
// all column loads in flight.
template <class T>
using global_ptr = __attribute__((address_space(1))) const T *;

__device__ __forceinline__ v4i load_stream4(const int32_t *p) {
    return __builtin_nontemporal_load((global_ptr<v4i>)p);
}

__device__ __forceinline__ int32_t load_stream1(const int32_t *p) {
    return __builtin_nontemporal_load((global_ptr<int32_t>)p);
}

// Relaxed device-scope atomic add on global memory (atomicAdd's ordering),
// issued as a global atomic for the reason given above.
__device__ __forceinline__ unsigned long long atomic_add_global(unsigned long long *p, unsigned long long v) {
    return __hip_atomic_fetch_add((__attribute__((address_space(1))) unsigned long long *)p, v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long atomic_xor_global(unsigned long long *p, unsigned long long v) {
    return __hip_atomic_fetch_xor((__attribute__((address_space(1))) unsigned long long *)p, v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

// Returns the value found at p (atomicCAS's contract).
__device__ __forceinline__ unsigned long long atomic_cas_global(unsigned long long *p, unsigned long long expect,
                                                                unsigned long long v) {
    __hip_atomic_compare_exchange_strong((__attribute__((address_space(1))) unsigned long long *)p, &expect, v,
                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return expect;
}

// Atomic adds on LDS through a pointer the compiler cannot place (an
// operator's shared-memory block is reached through a pointer member): the
// cast makes them ds_add instead of flat atomics.
__device__ __forceinline__ void atomic_add_lds(uint32_t *p, uint32_t v) {
    __hip_atomic_fetch_add((__attribute__((address_space(3))) uint32_t *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void atomic_add_lds(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_add((__attribute__((address_space(3))) unsigned long long *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Cached load of device memory (hash-table buckets) as a global load.
template <class T>
__device__ __forceinline__ T load_global(const T *p) {
    return *(global_ptr<T>)p;
}

// ------------------------------------------------------------------ hashing

// MurmurHash3 fmix32 finaliser; same function as hashMurmur (hashjoin.cu:4-11).
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// The inverse of fmix32 (each step undone in reverse order: x ^= x >> 16 is
// its own inverse, x ^= x >> 13 is undone by x ^= x >> 13 ^ x >> 26, and
// the multipliers by their inverses mod 2^32, 0xa5cb9243 * 0x85ebca6b =
// 0x7ed1b41d * 0xc2b2ae35 = 1): recovers a key from its hash, so a hashed
// row compares raw keys instead of hashing every slot.
__host__ __device__ __forceinline__ uint32_t fmix32_inv(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7ed1b41du;
    x ^= (x >> 13) ^ (x >> 26);
    x *= 0xa5cb9243u;
    x ^= x >> 16;
    return x;
}

// splitmix64 finaliser, used for order-independent join-pair digests.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// ------------------------------------------------------------------ Philox

struct u32x4 {
    uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds, key bumped between rounds.
__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
    }
    return c;
}

// Four consecutive words [4*blk, 4*blk+4) of stream (seed, stream).
__host__ __device__ __forceinline__ u32x4 philox_block(uint64_t seed, uint32_t stream, uint64_t blk) {
    u32x4 c = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__host__ __device__ __forceinline__ uint32_t philox_word(uint64_t seed, uint32_t stream, uint64_t i) {
    u32x4 r = philox_block(seed, stream, i >> 2);
    switch (i & 3) {
        case 0: return r.x;
        case 1: return r.y;
        case 2: return r.z;
        default: return r.w;
    }
}

// ------------------------------------------------------------------ Feistel

__host__ __device__ __forceinline__ uint32_t feistel_round_key(uint64_t seed, uint32_t stream, uint32_t r) {
    uint32_t inner = (uint32_t)(seed >> 32) + stream * 0x85EBCA6Bu + r * 0x9E3779B9u;
    return fmix32((uint32_t)seed ^ fmix32(inner));
}

// Keyed 4-round balanced Feistel bijection on [0, 2^nbits), nbits even.
__host__ __device__ __forceinline__ uint32_t feistel(uint32_t x, uint32_t nbits, uint64_t seed, uint32_t stream) {
    uint32_t h = nbits / 2;
    uint32_t m = (h >= 32) ? 0xffffffffu : ((1u << h) - 1u);
    uint32_t L = x >> h, R = x & m;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
        uint32_t F = fmix32(R ^ feistel_round_key(seed, stream, r)) & m;
        uint32_t nL = R;
        R = L ^ F;
        L = nL;
    }
    return (L << h) | R;
}

__host__ __device__ __forceinline__ uint32_t perm_width(uint32_t range) {
    uint32_t nb = 2;
    while (nb < 32 && ((uint64_t)1 << nb) < range) nb += 2;
    return nb;
}

// Bijection on [0, range): cycle-walk the Feistel network of the smallest even
// width covering `range` until the value falls inside the range.
__host__ __device__ __forceinline__ uint32_t perm(uint32_t x, uint32_t range, uint32_t nb, uint64_t seed,
                                                  uint32_t stream) {
    uint32_t y = x;
    do {
        y = feistel(y, nb, seed, stream);
    } while (y >= range);
    return y;
}

}  // namespace mgv
