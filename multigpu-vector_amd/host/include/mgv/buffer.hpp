// buffer.hpp -- buffer<T, cap>: a vector of up to `cap` rows of one or more
// int32 columns, in device or pinned host memory.  Counterpart of
// src/buffer.cuh:12-219.
//
// The reference's buffer object lives in device memory and is written by
// kernels (try_write / try_partial_final_write, buffer.cuh:92-134); here the
// object is a host-side handle and the device-side append is done inside
// libmgv's kernels (mgv_select_to_buffer, OpMaterialize), which reserve
// exactly the valid rows with one atomic per wave.  Columns beyond the first
// are an extension of the single-column reference buffer: the generator can
// slice several aligned columns into one vector.
#pragma once

#include <algorithm>

#include "common.hpp"

template <typename T = int32_t, uint32_t size = DEFAULT_BUFF_CAP>
class buffer_inspector;

template <typename T = int32_t, uint32_t size = DEFAULT_BUFF_CAP>
class buffer {
    static_assert(sizeof(T) == sizeof(int32_t), "Incompatible size");
    static_assert((size % (4 * WARPSIZE)) == 0, "Buffer must have a size multiple of 4*warpSize");

public:
    typedef buffer_inspector<T, size> inspector_t;
    static constexpr int kMaxCols = 4;

    uint32_t cnt = 0;                    // rows
    T *data = nullptr;                   // first column
    T *cols[kMaxCols] = {nullptr, nullptr, nullptr, nullptr};   // all columns (cols[0] == data)
    int ncols = 1;
    int device = -1;                     // >= 0 device memory, < 0 host memory
    int32_t pool_index = -1;             // index in the owning pool, -1 if not pooled
    bool owned = false;                  // pool buffers own their memory
    // A one-row scalar result (aggregation / join count) delivered to the
    // host: the row holds the low 32 bits the reference's int32 result
    // carries, value64 the exact value.
    bool scalar = false;
    int64_t value64 = 0;

    buffer() = default;
    buffer(T *d, int dev) : data(d), device(dev) { cols[0] = d; }

    uint32_t count() const { return std::min(cnt, size); }
    const T *get_data() const { return data; }
    uint32_t remaining_capacity() const { return size - count(); }
    static constexpr uint32_t capacity() { return size; }
    void clean() { cnt = 0; }
    bool full() const { return cnt >= size; }
    bool empty() const { return cnt == 0; }
    bool may_write() const { return cnt <= size - 4 * WARPSIZE; }

    void set_columns(T *const *c, int n) {
        ncols = std::min(n, kMaxCols);
        for (int i = 0; i < kMaxCols; ++i) cols[i] = i < ncols ? c[i] : nullptr;
        data = cols[0];
    }
};

// Host-side copies of a buffer's rows (buffer.cuh:165-218).
template <typename T, uint32_t size>
class buffer_inspector {
public:
    typedef buffer<T, size> buffer_t;
    buffer_t *buff = nullptr;
    mgv_stream_t strm;

    explicit buffer_inspector(mgv_stream_t s) : strm(s) {}

    void load(buffer_t *b, bool blocking = true) {
        buff = b;
        if (blocking) gpu(mgv_stream_synchronize(strm));
    }
    uint32_t count() const { return buff->count(); }
    const T *data() const { return buff->data; }
    static constexpr uint32_t capacity() { return buffer_t::capacity(); }

    void overwrite(const T *src, uint32_t N, bool blocking = true) {
        if (N > size) throw mgv::error(MGV_E_CAPACITY, "buffer_inspector::overwrite: too many rows");
        gpu(mgv_memcpy_async(buff->data, src, (size_t)N * sizeof(T), strm));
        buff->cnt = N;
        if (blocking) gpu(mgv_stream_synchronize(strm));
    }
    void read(T *dst, uint32_t N, bool blocking = true) {
        if (N > buff->cnt) throw mgv::error(MGV_E_INVALID, "buffer_inspector::read: too many rows");
        gpu(mgv_memcpy_async(dst, buff->data, (size_t)N * sizeof(T), strm));
        if (blocking) gpu(mgv_stream_synchronize(strm));
    }
    void save(buffer_t *, bool blocking = true) {
        if (blocking) gpu(mgv_stream_synchronize(strm));
    }
};
