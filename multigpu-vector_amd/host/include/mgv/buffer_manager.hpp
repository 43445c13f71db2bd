// buffer_manager.hpp -- buffer_manager<int32_t>: the process-wide buffer
// pools.  Counterpart of src/buffer_manager.cuh:40-181 / .cu:38-218.
//
// init(size, buff_buffer_size, keep_threshold) creates `size` device buffers
// (libmgv pool) and `size` pinned host buffers.  h_get_buffer(dev) hands out
// device buffers from a host-side cache refilled buff_buffer_size at a time;
// release_buffer returns a buffer to its pool (device buffers through the
// cache, batched back once it holds more than keep_threshold).  get_buffer()
// blocks until a host buffer is free.  Only device 0 / the current device is
// managed: this build runs on one GPU.
#pragma once

#include <condition_variable>
#include <memory>
#include <mutex>
#include <vector>

#include "buffer_pool.hpp"

template <typename T = int32_t>
class buffer_manager {
    static_assert(std::is_same<T, int32_t>::value, "Not implemented yet");

public:
    typedef buffer<int32_t, DEFAULT_BUFF_CAP> buffer_t;
    typedef buffer_pool<int32_t, DEFAULT_BUFF_CAP> pool_t;

    static void init(int size = 64, int buff_buffer_size = 8, int buff_keep_threshold = 16);
    static void destroy();
    static bool initialized();

    // host (pinned) buffers
    static buffer_t *get_buffer();
    static bool try_get_buffer2(buffer_t **ret);
    // device buffers
    static buffer_t *h_get_buffer(int dev);
    // any buffer handed out by this manager
    static void release_buffer(buffer_t *buff);
    // true for buffers handed out by get_buffer / h_get_buffer
    static bool owns(const buffer_t *buff);
    // H2D copy into a buffer (buffer_manager.cu:214-218)
    static void overwrite(buffer_t *buff, const T *data, uint32_t N, mgv_stream_t strm, bool blocking = true);

    static pool_t *device_pool();
    static uint32_t host_available();
    static uint32_t device_cached();

private:
    struct state;
    static state &S();
};
