// buffer_pool.hpp -- buffer_pool<T>: device buffers behind libmgv's
// lock-free device pool.  Counterpart of src/buffer_pool.cuh:35-170 with the
// same acquire / release / producer contract:
//   acquire_buffer_blocked_unsafe() -> a buffer, nullptr at end of stream
//   (no producer registered and no buffer available), or (buffer_t *)1 when
//   100 attempts failed (buffer_pool.cuh:83-96).
#pragma once

#include <vector>

#include "buffer.hpp"

template <typename T = int32_t, uint32_t buffer_capacity = DEFAULT_BUFF_CAP>
class buffer_pool {
public:
    typedef buffer<T, buffer_capacity> buffer_t;

private:
    mgv_pool *pool_ = nullptr;
    std::vector<buffer_t> bufs_;
    std::vector<uint32_t *> dev_cnt_;

public:
    buffer_pool(size_t size, size_t fill_size, int device = 0) {
        if (fill_size == 0 || fill_size > size)
            throw mgv::error(MGV_E_INVALID, "buffer_pool: need 0 < fill_size <= size");
        gpu(mgv_pool_create((uint32_t)fill_size, buffer_capacity, &pool_));
        bufs_.resize(fill_size);
        dev_cnt_.resize(fill_size);
        for (size_t i = 0; i < fill_size; ++i) {
            int32_t *d = nullptr;
            gpu(mgv_pool_buffer(pool_, (int32_t)i, &d, &dev_cnt_[i]));
            bufs_[i] = buffer_t((T *)d, device);
            bufs_[i].pool_index = (int32_t)i;
            bufs_[i].owned = true;
        }
    }

    ~buffer_pool() { mgv_pool_destroy(pool_); }
    buffer_pool(const buffer_pool &) = delete;
    buffer_pool &operator=(const buffer_pool &) = delete;

    static buffer_t *retry_sentinel() { return (buffer_t *)1; }
    static bool is_valid(buffer_t *x) { return x != nullptr; }

    buffer_t *acquire_buffer_blocked_unsafe(mgv_stream_t s = nullptr) {
        int32_t b = 0;
        gpu(mgv_pool_acquire(pool_, &b, s));
        if (b == MGV_POOL_END) return nullptr;
        if (b == MGV_POOL_RETRY) return retry_sentinel();
        bufs_[b].cnt = 0;
        bufs_[b].data = bufs_[b].cols[0];
        bufs_[b].ncols = 1;
        return &bufs_[b];
    }

    buffer_t *h_acquire_buffer_blocked(buffer_t **ret, mgv_stream_t s) {
        *ret = acquire_buffer_blocked_unsafe(s);
        return *ret;
    }

    void release_buffer(buffer_t *b, mgv_stream_t s = nullptr) {
        if (!b || b->pool_index < 0 || (size_t)b->pool_index >= bufs_.size() || b != &bufs_[b->pool_index])
            throw mgv::error(MGV_E_INVALID, "buffer_pool::release_buffer: not a buffer of this pool");
        gpu(mgv_pool_release(pool_, b->pool_index, s));
    }
    void h_release_buffer(buffer_t *b, mgv_stream_t s) { release_buffer(b, s); }

    void register_producer(void *, mgv_stream_t s = nullptr) {
        gpu(mgv_pool_register_producer(pool_, s));
        gpu(mgv_stream_synchronize(s));
    }
    void unregister_producer(void *, mgv_stream_t s = nullptr) {
        gpu(mgv_pool_unregister_producer(pool_, s));
        gpu(mgv_stream_synchronize(s));
    }

    // Row count written on the device (e.g. by mgv_select_to_buffer).
    uint32_t device_count(buffer_t *b) {
        uint32_t c = 0;
        gpu(mgv_memcpy(&c, dev_cnt_[b->pool_index], sizeof(c)));
        b->cnt = c;
        return b->count();
    }

    size_t size() const { return bufs_.size(); }
    mgv_pool *handle() const { return pool_; }

    void stats(uint32_t *available, uint32_t *producers, uint32_t *stacked) {
        gpu(mgv_pool_stats(pool_, available, producers, stacked, nullptr));
    }
};
