// buffer_manager.cpp -- process-wide pools of the C++ drop-in layer
// (counterpart of src/buffer_manager.cu:38-218).
#include "mgv/buffer_manager.hpp"

#include <chrono>
#include <thread>

template <>
struct buffer_manager<int32_t>::state {
    bool init = false;
    std::unique_ptr<pool_t> dpool;
    std::vector<buffer_t *> dcache;
    std::mutex dmu;
    int batch = 8, keep = 16;
    std::vector<buffer_t> hbufs;
    std::vector<int32_t *> hmem;
    std::vector<buffer_t *> hfree;
    std::mutex hmu;
    std::condition_variable hcv;
};

template <>
typename buffer_manager<int32_t>::state &buffer_manager<int32_t>::S() {
    static state s;
    return s;
}

template <>
void buffer_manager<int32_t>::init(int size, int buff_buffer_size, int buff_keep_threshold) {
    state &s = S();
    if (s.init) return;
    if (size < 1 || buff_buffer_size < 1 || buff_keep_threshold < buff_buffer_size)
        throw mgv::error(MGV_E_INVALID, "buffer_manager::init: bad sizes");
    gpu(mgv_init(0));
    s.dpool.reset(new pool_t((size_t)size, (size_t)size, 0));
    s.dpool->register_producer(nullptr);   // the manager produces buffers for the whole process
    s.batch = buff_buffer_size;
    s.keep = buff_keep_threshold;
    s.hbufs.resize(size);
    s.hmem.resize(size, nullptr);
    for (int i = 0; i < size; ++i) {
        void *p = nullptr;
        gpu(mgv_host_malloc(&p, (size_t)buffer_t::capacity() * sizeof(int32_t)));
        s.hmem[i] = (int32_t *)p;
        s.hbufs[i] = buffer_t(s.hmem[i], -1);
        s.hfree.push_back(&s.hbufs[i]);
    }
    s.init = true;
}

template <>
bool buffer_manager<int32_t>::initialized() {
    return S().init;
}

template <>
void buffer_manager<int32_t>::destroy() {
    state &s = S();
    if (!s.init) return;
    {
        std::lock_guard<std::mutex> lk(s.dmu);
        for (buffer_t *b : s.dcache) s.dpool->release_buffer(b);
        s.dcache.clear();
    }
    mgv_device_synchronize();
    s.dpool->unregister_producer(nullptr);
    s.dpool.reset();
    for (int32_t *p : s.hmem) mgv_host_free(p);
    s.hmem.clear();
    s.hbufs.clear();
    s.hfree.clear();
    s.init = false;
}

template <>
typename buffer_manager<int32_t>::buffer_t *buffer_manager<int32_t>::get_buffer() {
    state &s = S();
    if (!s.init) throw mgv::error(MGV_E_NOT_INIT, "buffer_manager::init was not called");
    std::unique_lock<std::mutex> lk(s.hmu);
    s.hcv.wait(lk, [&] { return !s.hfree.empty(); });
    buffer_t *b = s.hfree.back();
    s.hfree.pop_back();
    const size_t i = (size_t)(b - s.hbufs.data());
    *b = buffer_t(s.hmem[i], -1);
    return b;
}

template <>
bool buffer_manager<int32_t>::try_get_buffer2(buffer_t **ret) {
    state &s = S();
    std::lock_guard<std::mutex> lk(s.hmu);
    if (s.hfree.empty()) return false;
    buffer_t *b = s.hfree.back();
    s.hfree.pop_back();
    const size_t i = (size_t)(b - s.hbufs.data());
    *b = buffer_t(s.hmem[i], -1);
    *ret = b;
    return true;
}

template <>
typename buffer_manager<int32_t>::buffer_t *buffer_manager<int32_t>::h_get_buffer(int dev) {
    state &s = S();
    if (!s.init) throw mgv::error(MGV_E_NOT_INIT, "buffer_manager::init was not called");
    if (dev < 0) return get_buffer();
    std::lock_guard<std::mutex> lk(s.dmu);
    if (s.dcache.empty()) {
        for (int k = 0; k < s.batch; ++k) {
            buffer_t *b = nullptr;
            for (int attempt = 0; attempt < 1000; ++attempt) {
                b = s.dpool->acquire_buffer_blocked_unsafe();
                if (b != pool_t::retry_sentinel()) break;
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            if (!pool_t::is_valid(b) || b == pool_t::retry_sentinel()) break;
            s.dcache.push_back(b);
        }
        if (s.dcache.empty()) throw mgv::error(MGV_E_CAPACITY, "buffer_manager: device buffers exhausted");
    }
    buffer_t *b = s.dcache.back();
    s.dcache.pop_back();
    b->cnt = 0;
    b->set_columns(&b->cols[0], 1);
    return b;
}

template <>
void buffer_manager<int32_t>::release_buffer(buffer_t *buff) {
    state &s = S();
    if (!buff) return;
    if (!s.hbufs.empty() && buff >= s.hbufs.data() && buff < s.hbufs.data() + s.hbufs.size()) {
        std::lock_guard<std::mutex> lk(s.hmu);
        s.hfree.push_back(buff);
        s.hcv.notify_one();
        return;
    }
    if (buff->pool_index >= 0 && buff->device >= 0) {
        std::lock_guard<std::mutex> lk(s.dmu);
        s.dcache.push_back(buff);
        if ((int)s.dcache.size() > s.keep) {
            for (int k = 0; k < s.batch; ++k) {
                s.dpool->release_buffer(s.dcache.back());
                s.dcache.pop_back();
            }
        }
        return;
    }
    throw mgv::error(MGV_E_INVALID, "buffer_manager::release_buffer: buffer not owned by the manager");
}

template <>
bool buffer_manager<int32_t>::owns(const buffer_t *buff) {
    state &s = S();
    if (!buff || !s.init) return false;
    if (!s.hbufs.empty() && buff >= s.hbufs.data() && buff < s.hbufs.data() + s.hbufs.size()) return true;
    return buff->pool_index >= 0 && buff->device >= 0;
}

template <>
void buffer_manager<int32_t>::overwrite(buffer_t *buff, const int32_t *data, uint32_t N, mgv_stream_t strm,
                                        bool blocking) {
    if (N > buffer_t::capacity()) throw mgv::error(MGV_E_CAPACITY, "buffer_manager::overwrite: too many rows");
    gpu(mgv_memcpy_async(buff->data, data, (size_t)N * sizeof(int32_t), strm));
    buff->cnt = N;
    if (blocking) gpu(mgv_stream_synchronize(strm));
}

template <>
typename buffer_manager<int32_t>::pool_t *buffer_manager<int32_t>::device_pool() {
    return S().dpool.get();
}

template <>
uint32_t buffer_manager<int32_t>::host_available() {
    std::lock_guard<std::mutex> lk(S().hmu);
    return (uint32_t)S().hfree.size();
}

template <>
uint32_t buffer_manager<int32_t>::device_cached() {
    std::lock_guard<std::mutex> lk(S().dmu);
    return (uint32_t)S().dcache.size();
}
