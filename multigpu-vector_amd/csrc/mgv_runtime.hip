// mgv_runtime.hip -- runtime part of the C ABI (device, memory, streams,
// events, graphs), synthetic-data generators and fills.
//
// Replaces the reference's gpu() error macro that calls exit()
// (src/common.cuh:12-23) with status codes, and the device side of
// buffer_manager<int32_t>::init (src/buffer_manager.cu:38-211) with mgv_init.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "mgv_internal.hpp"
#include "mgv_wave.hpp"

namespace mgv {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static std::mutex g_state_mu;
static DeviceState g_state[64];

DeviceState *current_state() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    DeviceState *st = &g_state[dev];
    return st->init ? st : nullptr;
}

// ------------------------------------------------------------------ kernels

__global__ void k_fill_u32(uint32_t *dst, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

__global__ void k_fill_u64(uint64_t *dst, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// One Philox block (4 words) per thread iteration; dst written as int4 when
// the 4 elements are in range and aligned.
template <typename F>
__global__ void k_gen_words(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, F f) {
    const uint64_t nblk = (n + 3) / 4;
    const bool vec = ((uintptr_t)dst & 15u) == 0;
    for (uint64_t blk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; blk < nblk;
         blk += (uint64_t)gridDim.x * blockDim.x) {
        u32x4 r = philox_block(seed, stream, blk);
        uint64_t i = blk * 4;
        if (vec && i + 3 < n) {
            v4i o = {f(r.x), f(r.y), f(r.z), f(r.w)};
            *reinterpret_cast<v4i *>(dst + i) = o;
        } else {
            uint32_t w[4] = {r.x, r.y, r.z, r.w};
            for (int j = 0; j < 4; ++j)
                if (i + j < n) dst[i + j] = f(w[j]);
        }
    }
}

struct FBits {
    uint32_t bits;
    __device__ int32_t operator()(uint32_t w) const { return bits ? (int32_t)(w >> (32 - bits)) : 0; }
};

struct FRange {
    uint32_t range;
    __device__ int32_t operator()(uint32_t w) const { return (int32_t)(uint32_t)(((uint64_t)w * range) >> 32); }
};

__global__ void k_gen_perm(int32_t *dst, uint64_t n, uint32_t range, uint32_t nb, uint64_t seed, uint32_t pstream) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = (int32_t)perm((uint32_t)i, range, nb, seed, pstream);
}

__global__ void k_gen_iota(int32_t *dst, uint64_t n, int32_t base) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = base + (int32_t)(uint32_t)i;
}

__device__ __forceinline__ uint32_t zipf_rank(uint32_t u, const uint64_t *T, uint32_t n) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)u < T[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__global__ void k_gen_zipf(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, const uint64_t *T,
                           uint32_t nranks, uint32_t nb, uint64_t pseed, uint32_t pstream) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t r = zipf_rank(philox_word(seed, stream, i), T, nranks);
        dst[i] = (int32_t)(pseed ? perm(r, nranks, nb, pseed, pstream) : r);
    }
}

static unsigned grid_for(uint64_t work, unsigned block) {
    uint64_t g = (work + block - 1) / block;
    DeviceState *st = current_state();
    uint64_t cap = (uint64_t)((st && st->cus) ? st->cus : 256) * 32;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace mgv

using namespace mgv;

#define MGV_REQUIRE_INIT()                                                            \
    do {                                                                              \
        if (!current_state()) return fail(MGV_E_NOT_INIT, "mgv_init was not called"); \
    } while (0)

extern "C" {

int mgv_abi_version(void) { return MGV_ABI_VERSION; }

const char *mgv_last_error_string(void) { return g_err; }

int mgv_init(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(MGV_E_NOGPU, "no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count || device >= 64) return fail(MGV_E_INVALID, "device %d out of range", device);
    int rc = hip_status(hipSetDevice(device), "hipSetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_state_mu);
    DeviceState &st = g_state[device];
    if (st.init) return MGV_OK;
    hipDeviceProp_t p;
    rc = hip_status(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (rc) return rc;
    st.device = device;
    st.cus = p.multiProcessorCount;
    st.lds_per_block = p.sharedMemPerBlock;
    rc = hip_status(hipStreamCreateWithFlags(&st.default_stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) return rc;
    if ((rc = probe_scratch_setup(st))) {
        (void)hipStreamDestroy(st.default_stream);
        st = DeviceState();
        return rc;
    }
    st.init = true;
    return MGV_OK;
}

int mgv_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_state_mu);
    for (auto &st : g_state) {
        if (!st.init) continue;
        (void)hipSetDevice(st.device);
        (void)hipStreamSynchronize(st.default_stream);
        (void)hipStreamDestroy(st.default_stream);
        st = DeviceState();
    }
    return MGV_OK;
}

int mgv_get_device_props(mgv_device_props *out) {
    MGV_REQUIRE_INIT();
    if (!out) return fail(MGV_E_INVALID, "null props");
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    int rc = hip_status(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    if (rc) return rc;
    memset(out, 0, sizeof(*out));
    strncpy(out->name, p.name, sizeof(out->name) - 1);
    strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->cus = p.multiProcessorCount;
    out->wave_size = p.warpSize;
    out->lds_per_block = (int64_t)p.sharedMemPerBlock;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    out->l2_bytes = (int64_t)p.l2CacheSize;
    out->clock_khz = p.clockRate;
    out->mem_clock_khz = p.memoryClockRate;
    out->mem_bus_width = p.memoryBusWidth;
    out->max_threads_per_block = p.maxThreadsPerBlock;
    return MGV_OK;
}

int mgv_malloc(void **ptr, size_t bytes) {
    MGV_REQUIRE_INIT();
    if (!ptr) return fail(MGV_E_INVALID, "null ptr");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    return hip_status(hipMalloc(ptr, bytes), "hipMalloc");
}

int mgv_free(void *ptr) {
    if (!ptr) return MGV_OK;
    return hip_status(hipFree(ptr), "hipFree");
}

int mgv_host_malloc(void **ptr, size_t bytes) {
    MGV_REQUIRE_INIT();
    if (!ptr) return fail(MGV_E_INVALID, "null ptr");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    return hip_status(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}

int mgv_host_free(void *ptr) {
    if (!ptr) return MGV_OK;
    return hip_status(hipHostFree(ptr), "hipHostFree");
}

int mgv_mem_get_info(size_t *free_bytes, size_t *total_bytes) {
    MGV_REQUIRE_INIT();
    return hip_status(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int mgv_memcpy(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MGV_OK;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
}

int mgv_memcpy_async(void *dst, const void *src, size_t bytes, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (bytes == 0) return MGV_OK;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream_of(s)), "hipMemcpyAsync");
}

int mgv_fill_u32(uint32_t *dst, uint64_t n, uint32_t value, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst) return fail(MGV_E_INVALID, "null dst");
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, stream_of(s), dst, n, value);
    return check_launch("k_fill_u32");
}

int mgv_fill_u64(uint64_t *dst, uint64_t n, uint64_t value, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst) return fail(MGV_E_INVALID, "null dst");
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, stream_of(s), dst, n, value);
    return check_launch("k_fill_u64");
}

int mgv_stream_create(mgv_stream_t *s) {
    MGV_REQUIRE_INIT();
    hipStream_t h;
    int rc = hip_status(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) *s = (mgv_stream_t)h;
    return rc;
}

int mgv_stream_destroy(mgv_stream_t s) {
    if (!s) return MGV_OK;
    return hip_status(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
}

int mgv_stream_synchronize(mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    return hip_status(hipStreamSynchronize(stream_of(s)), "hipStreamSynchronize");
}

int mgv_device_synchronize(void) { return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int mgv_event_create(mgv_event_t *e) {
    MGV_REQUIRE_INIT();
    hipEvent_t h;
    int rc = hip_status(hipEventCreate(&h), "hipEventCreate");
    if (!rc) *e = (mgv_event_t)h;
    return rc;
}

int mgv_event_destroy(mgv_event_t e) {
    if (!e) return MGV_OK;
    return hip_status(hipEventDestroy((hipEvent_t)e), "hipEventDestroy");
}

int mgv_event_record(mgv_event_t e, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    return hip_status(hipEventRecord((hipEvent_t)e, stream_of(s)), "hipEventRecord");
}

int mgv_event_synchronize(mgv_event_t e) {
    return hip_status(hipEventSynchronize((hipEvent_t)e), "hipEventSynchronize");
}

int mgv_event_query(mgv_event_t e) {
    hipError_t r = hipEventQuery((hipEvent_t)e);
    if (r == hipSuccess) return 0;
    if (r == hipErrorNotReady) return 1;
    return hip_status(r, "hipEventQuery");
}

int mgv_event_elapsed_ms(float *ms, mgv_event_t start, mgv_event_t stop) {
    return hip_status(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

int mgv_stream_wait_event(mgv_stream_t s, mgv_event_t e) {
    MGV_REQUIRE_INIT();
    return hip_status(hipStreamWaitEvent(stream_of(s), (hipEvent_t)e, 0), "hipStreamWaitEvent");
}

struct GraphBox {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int mgv_graph_begin_capture(mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    return hip_status(hipStreamBeginCapture(stream_of(s), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

int mgv_graph_end_capture(mgv_stream_t s, mgv_graph_t *g) {
    MGV_REQUIRE_INIT();
    hipGraph_t graph = nullptr;
    int rc = hip_status(hipStreamEndCapture(stream_of(s), &graph), "hipStreamEndCapture");
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    rc = hip_status(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) {
        (void)hipGraphDestroy(graph);
        return rc;
    }
    GraphBox *b = new GraphBox{graph, exec};
    *g = (mgv_graph_t)b;
    return MGV_OK;
}

int mgv_graph_launch(mgv_graph_t g, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!g) return fail(MGV_E_INVALID, "null graph");
    return hip_status(hipGraphLaunch(((GraphBox *)g)->exec, stream_of(s)), "hipGraphLaunch");
}

int mgv_graph_destroy(mgv_graph_t g) {
    if (!g) return MGV_OK;
    GraphBox *b = (GraphBox *)g;
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    delete b;
    return MGV_OK;
}

// ------------------------------------------------------------------ datagen

int mgv_gen_bits(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t bits, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst || bits > 32) return fail(MGV_E_INVALID, "mgv_gen_bits: bad args");
    hipLaunchKernelGGL(k_gen_words<FBits>, dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, stream_of(s), dst, n, seed,
                       stream, FBits{bits});
    return check_launch("k_gen_bits");
}

int mgv_gen_range(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t range, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst) return fail(MGV_E_INVALID, "mgv_gen_range: bad args");
    hipLaunchKernelGGL(k_gen_words<FRange>, dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, stream_of(s), dst, n,
                       seed, stream, FRange{range});
    return check_launch("k_gen_range");
}

int mgv_gen_perm(int32_t *dst, uint64_t n, uint32_t range, uint64_t seed, uint32_t pstream, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst || range == 0 || n > range) return fail(MGV_E_INVALID, "mgv_gen_perm: need 0 < n <= range");
    hipLaunchKernelGGL(k_gen_perm, dim3(grid_for(n, 256)), dim3(256), 0, stream_of(s), dst, n, range,
                       perm_width(range), seed, pstream);
    return check_launch("k_gen_perm");
}

int mgv_gen_iota(int32_t *dst, uint64_t n, int32_t base, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst) return fail(MGV_E_INVALID, "mgv_gen_iota: bad args");
    hipLaunchKernelGGL(k_gen_iota, dim3(grid_for(n, 256)), dim3(256), 0, stream_of(s), dst, n, base);
    return check_launch("k_gen_iota");
}

int mgv_zipf_table_create(uint64_t **table, uint32_t nranks, double s_exp) {
    MGV_REQUIRE_INIT();
    if (!table || nranks == 0) return fail(MGV_E_INVALID, "mgv_zipf_table_create: bad args");
    std::vector<uint64_t> T(nranks);
    double H = 0.0;
    for (uint32_t r = 0; r < nranks; ++r) H += (s_exp == 1.0) ? 1.0 / (double)(r + 1) : 1.0 / pow((double)(r + 1), s_exp);
    double acc = 0.0;
    for (uint32_t r = 0; r < nranks; ++r) {
        acc += (s_exp == 1.0) ? 1.0 / (double)(r + 1) : 1.0 / pow((double)(r + 1), s_exp);
        double f = acc / H * 4294967296.0;
        T[r] = (f >= 4294967296.0) ? 4294967296ull : (uint64_t)f;
    }
    T[nranks - 1] = 4294967296ull;
    void *d = nullptr;
    int rc = hip_status(hipMalloc(&d, nranks * sizeof(uint64_t)), "hipMalloc");
    if (rc) return rc;
    rc = hip_status(hipMemcpy(d, T.data(), nranks * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) {
        (void)hipFree(d);
        return rc;
    }
    *table = (uint64_t *)d;
    return MGV_OK;
}

int mgv_gen_zipf(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, const uint64_t *table, uint32_t nranks,
                 uint64_t pseed, uint32_t pstream, mgv_stream_t s) {
    MGV_REQUIRE_INIT();
    if (!n) return MGV_OK;
    if (!dst || !table || nranks == 0) return fail(MGV_E_INVALID, "mgv_gen_zipf: bad args");
    hipLaunchKernelGGL(k_gen_zipf, dim3(grid_for(n, 256)), dim3(256), 0, stream_of(s), dst, n, seed, stream, table,
                       nranks, perm_width(nranks), pseed, pstream);
    return check_launch("k_gen_zipf");
}

}  // extern "C"
