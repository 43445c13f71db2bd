// operators.cpp -- operator framework of the C++ drop-in layer
// (counterparts: src/operators/operator.cu, exchange.cu, generators.cu,
// gpu_to_cpu.cu, materializer.cu, hashjoin.cu:63-75).
#include "mgv/operators.hpp"

#include <chrono>
#include <algorithm>
#include <climits>
#include <cstring>

// ------------------------------------------------------------------ hashjoin

template <>
hashjoin<WARPSIZE>::hashjoin(d_operator_t *parent_, int32_t log_table_size, int32_t max_size, const launch_conf &conf)
    : parent(parent_), table(std::make_shared<mgv_table_state>()) {
    if (max_size < 0) throw mgv::error(MGV_E_INVALID, "hashjoin: negative max_size");
    if (log_table_size < 0 || log_table_size > 31) throw mgv::error(MGV_E_INVALID, "hashjoin: log_table_size in [0, 31]");
    uint64_t bytes = 0;
    gpu(mgv_ht_plan((uint64_t)max_size, 0.5, &table->nbuckets, &bytes));
    // at least 2^log_table_size slots
    const uint64_t min_buckets = ((1ull << log_table_size) + MGV_SLOTS_PER_BUCKET - 1) / MGV_SLOTS_PER_BUCKET;
    if (min_buckets > table->nbuckets) {
        if (min_buckets > 0xffffffffull) throw mgv::error(MGV_E_CAPACITY, "hashjoin: table too large");
        table->nbuckets = (uint32_t)min_buckets;
        bytes = min_buckets * MGV_SLOTS_PER_BUCKET * sizeof(uint64_t);
    }
    void *p = nullptr;
    gpu(mgv_malloc(&p, bytes));
    table->slots = (uint64_t *)p;
    table->max_size = (uint64_t)max_size;
    gpu(mgv_ht_clear(table->slots, table->nbuckets, nullptr));
    gpu(mgv_stream_synchronize(nullptr));
    builder = d_operator_t::create<hashjoin_builder_t>(conf, table);
}

template <>
hashjoin<WARPSIZE>::~hashjoin() {
    d_operator_t::destroy(builder);
}

// ------------------------------------------------------------------ gpu_to_cpu

template <>
gpu_to_cpu<WARPSIZE, 64, buffer_t *>::gpu_to_cpu(h_operator_t *parent_, int device_) : parent(parent_), device(device_) {
    if (!buffer_manager<int32_t>::initialized())
        throw mgv::error(MGV_E_NOT_INIT, "gpu_to_cpu: buffer_manager<int32_t>::init was not called");
    try {
        gpu(mgv_pool_create_host(kOutBuffers, kOutRows, &out_pool));
        // a slot for every pool buffer, the end-of-stream entry and a scalar result
        gpu(mgv_ring_create(kOutBuffers + 4, &ring));
        gpu(mgv_stream_out_create(out_pool, ring, &rows_out));
    } catch (...) {
        mgv_ring_destroy(ring);
        mgv_pool_destroy(out_pool);
        throw;
    }
}

template <>
gpu_to_cpu<WARPSIZE, 64, buffer_t *>::~gpu_to_cpu() {
    stop = true;
    if (catcher.joinable()) catcher.join();
    mgv_stream_out_destroy(rows_out);
    mgv_ring_destroy(ring);
    mgv_pool_destroy(out_pool);
}

// ------------------------------------------------------------------ split / union_all

template <>
split<WARPSIZE, int32_t>::split(std::vector<d_operator_t *> parents_, launch_conf conf, int dev)
    : parents(std::move(parents_)) {
    (void)conf;
    (void)dev;
    if (parents.empty() || parents.size() > MGV_SPLIT_MAX_PARENTS)
        throw mgv::error(MGV_E_INVALID, "split: 1 to 8 parents");
    for (d_operator_t *p : parents)
        if (!p) throw mgv::error(MGV_E_INVALID, "split: null parent");
}

template <>
union_all<WARPSIZE, int32_t>::union_all(d_operator_t *parent_, int num_of_children_, launch_conf conf, int dev)
    : parent(parent_), num_of_children(num_of_children_) {
    (void)conf;
    (void)dev;
    if (!parent) throw mgv::error(MGV_E_INVALID, "union_all: null parent");
    if (num_of_children < 1) throw mgv::error(MGV_E_INVALID, "union_all: at least one child");
}

// ------------------------------------------------------------------ d_operator_t

namespace {

d_operator_t *parent_of(const d_operator_t::op_t &op) {
    if (auto p = std::get_if<select_t *>(&op)) return (*p)->parent;
    if (auto p = std::get_if<aggregation_t *>(&op)) return (*p)->parent;
    if (auto p = std::get_if<hashjoin_t *>(&op)) return (*p)->parent;
    if (auto p = std::get_if<union_all_t *>(&op)) return (*p)->parent;
    return nullptr;
}

}  // namespace

void d_operator_t::resolve() {
    // the chain from this operator up to its sink; a union_all passes its
    // children's vectors on to its parent (union_all.cu:45-48), so it is not
    // a stage of the chain: a union_all at the head owns the chain's
    // accumulators, one further up shares its own with this chain
    std::vector<d_operator_t *> chain;
    d_operator_t *start = this;
    if (std::holds_alternative<union_all_t *>(op)) start = parent_of(op);
    size_t hops = 0;
    for (d_operator_t *c = start; c; c = parent_of(c->op)) {
        if (++hops > 9) throw mgv::error(MGV_E_INVALID, "operator chain too long or cyclic");
        if (std::holds_alternative<union_all_t *>(c->op)) {
            if (through || c == this) throw mgv::error(MGV_E_INVALID, "at most one union_all per operator chain");
            through = c;
            continue;
        }
        chain.push_back(c);
    }
    auto is = [&](size_t i, auto tag) {
        using P = decltype(tag);
        return i < chain.size() && std::holds_alternative<P>(chain[i]->op);
    };
    size_t i = 0;
    bool sel = false;
    if (is(0, (select_t *)nullptr)) {
        sel = true;
        bound = std::get<select_t *>(chain[0]->op)->filt.bound;
        i = 1;
    }
    filtered = sel;
    if (is(i, (split_t *)nullptr)) {
        // [select ->] split: the split's parents are the heads of their own chains
        if (through) throw mgv::error(MGV_E_INVALID, "a split after a union_all must be its first operator");
        kind = SPLIT;
        splitter = std::get<split_t *>(chain[i]->op);
        return;
    }
    if (is(i, (hashjoin_builder_t *)nullptr) && chain.size() == i + 1) {
        kind = sel ? SELECT_BUILD : BUILD;
        table = std::get<hashjoin_builder_t *>(chain[i]->op)->table;
        return;
    }
    if (is(i, (hashjoin_t *)nullptr)) {
        table = std::get<hashjoin_t *>(chain[i]->op)->table;
        kind = sel ? SELECT_PROBE : PROBE;
        ++i;
        // hashjoin hands its match count to its parent once (hashjoin.cu:87-96):
        // an aggregation above it sums that single value, i.e. the count
        if (is(i, (aggregation_t *)nullptr)) ++i;
    } else if (is(i, (aggregation_t *)nullptr)) {
        kind = SELECT_SUM;
        if (!sel) bound = INT32_MAX;
        sum = true;
        ++i;
    } else if (sel) {
        kind = SELECT_MAT;
    } else if (is(0, (gpu_to_cpu_t *)nullptr)) {
        // gpu_to_cpu fed directly: every row goes to the host
        // (create_select_all_pipeline, main9.cu:176-200)
        kind = SELECT_MAT;
        bound = INT32_MAX;
    }
    if (kind == NONE || !is(i, (gpu_to_cpu_t *)nullptr) || i + 1 != chain.size())
        throw mgv::error(MGV_E_INVALID, "unsupported device operator chain");
    sink = std::get<gpu_to_cpu_t *>(chain[i]->op);
}

uint32_t d_operator_t::split_dests() const {
    const size_t np = splitter->parents.size();
    return np < 2 ? (uint32_t)np : 2u;   // x & 1 names parents 0 and 1 only
}

void d_operator_t::open() {
    if (auto u = std::get_if<union_all_t *>(&op)) {
        // the first child opens the parent (union_all.cu:19-30)
        std::lock_guard<std::mutex> g((*u)->host_lock);
        if ((*u)->num_of_active_children++ == 0) do_open();
        return;
    }
    do_open();
}

void d_operator_t::close() {
    if (auto u = std::get_if<union_all_t *>(&op)) {
        // the last child closes the parent (union_all.cu:63-74)
        std::lock_guard<std::mutex> g((*u)->host_lock);
        if (++(*u)->num_of_closed_children < (*u)->num_of_children) return;
        (*u)->num_of_active_children = 0;
        (*u)->num_of_closed_children = 0;
        do_close();
        return;
    }
    do_close();
}

void d_operator_t::do_open() {
    if (opened) return;
    if (kind == NONE) resolve();
    if (through) {
        // one child of the union_all: the first child's open opens the
        // shared chain (union_all.cu:19-30)
        through->open();
        opened = true;
        return;
    }
    if (kind == SPLIT) {
        if (!cursors) {
            void *p = nullptr;
            gpu(mgv_malloc(&p, MGV_SPLIT_MAX_PARENTS * sizeof(uint64_t)));
            cursors = (uint64_t *)p;
        }
        gpu(mgv_fill_u64(cursors, MGV_SPLIT_MAX_PARENTS, 0, nullptr));
        gpu(mgv_stream_synchronize(nullptr));
        stage_upper = 0;
        opened = true;
        // before_open opens every parent (split.cu:22-34)
        for (d_operator_t *p : splitter->parents) p->open();
        return;
    }
    if (sink) {
        // before_open: open the host parent and start the catcher
        sink->parent->open();
        sink->stop = false;
        sink->failure.clear();
        sink->last = mgv_result{};
        sink->buffers = sink->dropped = 0;
        gpu_to_cpu_t *s = sink;
        const bool rows = kind == SELECT_MAT;
        const bool scalar_sum = sum;
        if (rows) {
            gpu(mgv_stream_out_open(s->rows_out, nullptr));
            gpu(mgv_stream_synchronize(nullptr));
        }
        s->catcher = std::thread([s, rows, scalar_sum]() {
            try {
                mgv_ring_entry e;
                for (;;) {
                    const int rc = mgv_ring_poll(s->ring, &e);
                    if (rc == 1) {
                        if (s->stop.load()) break;
                        std::this_thread::sleep_for(std::chrono::microseconds(20));
                        continue;
                    }
                    gpu(rc);
                    if (!rows) {   // the scalar result (aggregation / join count)
                        s->last = e.res;
                        const int64_t v = scalar_sum ? e.res.sum : (int64_t)e.res.count;
                        s->rows.assign(1, (int32_t)(uint32_t)(uint64_t)v);
                        s->out.assign(1, buffer_t(s->rows.data(), -1));
                        buffer_t &b = s->out[0];
                        b.cnt = 1;
                        b.scalar = true;
                        b.value64 = v;
                        s->parent->consume(&b);
                        break;
                    }
                    if (e.id == MGV_RING_EOF) {   // at_close: totals of the stream
                        s->last = mgv_result{};
                        s->last.count = e.res.count;
                        s->dropped = e.res.overflow;
                        break;
                    }
                    // a full pool buffer (pinned host memory): copy its rows
                    // into a host buffer of the manager, hand the pool buffer
                    // back, pass the rows on
                    if (e.value > gpu_to_cpu_t::kOutRows) throw mgv::error(MGV_E_CAPACITY, "gpu_to_cpu: bad buffer");
                    int32_t *data = nullptr;
                    gpu(mgv_pool_buffer(s->out_pool, (int32_t)e.id, &data, nullptr));
                    buffer_t *h = buffer_manager<int32_t>::get_buffer();
                    memcpy(h->data, data, (size_t)e.value * sizeof(int32_t));
                    gpu(mgv_pool_release_host(s->out_pool, (int32_t)e.id));
                    h->cnt = (uint32_t)e.value;
                    ++s->buffers;
                    s->parent->consume(h);
                }
            } catch (const std::exception &ex) {
                s->failure = ex.what();
            }
        });
    }
    // at_open: reset the accumulators
    if (kind != BUILD && kind != SELECT_BUILD) {
        if (!res) {
            void *p = nullptr;
            gpu(mgv_malloc(&p, sizeof(mgv_result)));
            res = (mgv_result *)p;
        }
        gpu(mgv_result_clear(res, 1, nullptr));
    }
    gpu(mgv_stream_synchronize(nullptr));
    opened = true;
}

void d_operator_t::consume(buffer_t *v, mgv_stream_t s) {
    if (!opened) do_open();
    // the accumulators: this chain's own, or those of the union_all it
    // feeds; the children of a union_all consume one at a time
    d_operator_t *const acc = through ? through : this;
    std::unique_lock<std::mutex> fan_in;
    if (auto u = std::get_if<union_all_t *>(&acc->op)) fan_in = std::unique_lock<std::mutex>((*u)->host_lock);
    mgv_result *const res = acc->res;
    last_stream = s;
    const uint64_t n = v->count();
    if (n == 0) return;
    const mgv_launch L = conf.to_launch();
    int32_t *const *c = v->cols;
    const int nc = v->ncols;
    switch (kind) {
        case SELECT_SUM:
            gpu(mgv_select_sum(c[0], nc > 1 ? c[1] : c[0], n, bound, res, L, s));
            break;
        case SELECT_MAT:
            // gpu_to_cpu::consume_warp: the survivors stream into the sink's
            // rotating pool buffers (gpu_to_cpu.cu:61-105)
            gpu(mgv_select_stream(c[0], nc > 1 ? c[1] : c[0], n, bound, sink->rows_out, L, s));
            break;
        case PROBE:
        case SELECT_PROBE: {
            const int32_t *pred = kind == SELECT_PROBE ? c[0] : nullptr;
            const int32_t *keys = kind == SELECT_PROBE && nc > 1 ? c[1] : c[0];
            uint64_t need = 0;
            if (lds_join(n, false, &need)) {   // a large vector against a large table: through LDS slices
                void *w = workspace(s, need);
                gpu(mgv_hashjoin_probe_lds(table->slots, table->nbuckets, pred, bound, keys, nullptr, n, kLdsFanout, w,
                                           need, res, mgv_launch{0, 1024, -1, MGV_LAUNCH_SCATTER(2u)}, s));
            } else {
                gpu(mgv_hashjoin_probe(table->slots, table->nbuckets, pred, bound, keys, nullptr, n, 0, 0, res, L, s));
            }
            break;
        }
        case BUILD:
        case SELECT_BUILD: {
            const int32_t *pred = kind == SELECT_BUILD ? c[0] : nullptr;
            const int k = kind == SELECT_BUILD ? 1 : 0;   // first carried column: the key
            const int32_t *keys = nc > k ? c[k] : c[0];
            const int32_t *pay = nc > k + 1 ? c[k + 1] : nullptr;
            const int mode = pay ? MGV_PAYLOAD_COLUMN : MGV_PAYLOAD_NONE;
            uint64_t need = 0;
            if (lds_join(n, true, &need)) {
                void *w = workspace(s, need);
                gpu(mgv_hashjoin_build_lds(table->slots, table->nbuckets, keys, pay, mode, pred, bound, n, 0, kLdsFanout,
                                           w, need, nullptr, mgv_launch{0, 1024, -1, MGV_LAUNCH_SCATTER(2u)}, s));
            } else {
                gpu(mgv_hashjoin_build(table->slots, table->nbuckets, keys, pay, mode, pred, bound, n, 0, nullptr, L,
                                       s));
            }
            break;
        }
        case SPLIT: {
            // the select's filter column, then the columns the split carries
            // (the first decides the destination)
            const int32_t *pred = filtered ? c[0] : nullptr;
            int32_t *const *car = (filtered && nc > 1) ? c + 1 : c;
            const int ncar = (filtered && nc > 1) ? nc - 1 : nc;
            const uint32_t nd = split_dests();
            if (!stage || stage_cols != ncar || n > stage_cap) {
                flush_split(s);
                gpu(mgv_stream_synchronize(s));
                gpu(mgv_free(stage));
                stage = nullptr;
                const uint64_t cap = std::max<uint64_t>(2 * n, buffer_t::capacity());
                void *p = nullptr;
                gpu(mgv_malloc(&p, cap * nd * (uint64_t)ncar * sizeof(int32_t)));
                stage = (int32_t *)p;
                stage_cap = cap;
                stage_cols = ncar;
            }
            if (stage_upper + n > stage_cap) flush_split(s);
            int32_t *outs[MGV_SPLIT_MAX_PARENTS * MGV_SPLIT_MAX_COLS] = {};
            for (uint32_t d = 0; d < nd; ++d)
                for (int k = 0; k < ncar; ++k) outs[d * ncar + k] = stage + (d * (uint64_t)ncar + k) * stage_cap;
            gpu(mgv_split(pred, bound, (const int32_t *const *)car, (uint32_t)ncar, n, nd, outs, stage_cap, cursors,
                          L, s));
            stage_upper += n;   // an upper bound of every destination's staged rows
            break;
        }
        default:
            throw mgv::error(MGV_E_INVALID, "d_operator_t::consume: operator is not the head of a chain");
    }
}

// Vectors of at least kLdsRows rows against a table of at least kLdsTableBytes
// (far larger than the L2) are built and probed through LDS table slices
// (mgv_hashjoin_build_lds / mgv_hashjoin_probe_lds) when the table's buckets
// split into slices; the results are those of the direct kernels.
bool d_operator_t::lds_join(uint64_t n, bool build, uint64_t *need) const {
    if (n < (build ? kLdsBuildRows : kLdsRows) || (uint64_t)table->nbuckets * 64 < kLdsTableBytes) return false;
    return (build ? mgv_build_lds_workspace_size(n, table->nbuckets, kLdsFanout, need)
                  : mgv_probe_lds_workspace_size(n, table->nbuckets, kLdsFanout, need)) == MGV_OK;
}

// The chain's workspace for stream s, at least `bytes`: reused vector after
// vector in stream order, replaced (after the stream's work on it is done)
// when a larger one is needed.
void *d_operator_t::workspace(mgv_stream_t s, uint64_t bytes) {
    std::lock_guard<std::mutex> g(ws_mu);
    auto &w = ws[s];
    if (w.second < bytes) {
        if (w.first) {
            gpu(mgv_stream_synchronize(s));
            gpu(mgv_free(w.first));
            w = {nullptr, 0};
        }
        gpu(mgv_malloc(&w.first, bytes));
        w.second = bytes;
    }
    return w.first;
}

// Hands the staged rows of every destination to its parent as vectors of at
// most the buffer capacity, on stream s, then empties the staging (the
// reference's split forwards each full 4*32-row chunk at once, split.cu:95-107,
// and what is left at close, 131-150).
void d_operator_t::flush_split(mgv_stream_t s) {
    if (stage_upper == 0) return;
    gpu(mgv_stream_synchronize(s));
    uint64_t cnt[MGV_SPLIT_MAX_PARENTS] = {};
    const uint32_t nd = split_dests();
    gpu(mgv_memcpy(cnt, cursors, nd * sizeof(uint64_t)));
    const uint64_t cap = buffer_t::capacity();
    for (uint32_t d = 0; d < nd; ++d) {
        if (cnt[d] > stage_cap) throw mgv::error(MGV_E_CAPACITY, "split: staging overflow");
        for (uint64_t off = 0; off < cnt[d]; off += cap) {
            int32_t *cols[buffer_t::kMaxCols] = {};
            for (int k = 0; k < stage_cols; ++k) cols[k] = stage + (d * (uint64_t)stage_cols + k) * stage_cap + off;
            buffer_t b(cols[0], conf.device);
            b.set_columns(cols, stage_cols);
            b.cnt = (uint32_t)std::min<uint64_t>(cap, cnt[d] - off);
            splitter->parents[d]->consume(&b, s);
        }
    }
    // the parents' launches on s read the staging before this clear and the
    // next split launch
    gpu(mgv_fill_u64(cursors, nd, 0, s));
    stage_upper = 0;
}

void d_operator_t::do_close() {
    if (!opened) return;
    if (through) {
        // the last child's close closes the shared chain (union_all.cu:63-74)
        gpu(mgv_stream_synchronize(last_stream));
        opened = false;
        through->close();
        return;
    }
    if (kind == SPLIT) {
        flush_split(last_stream);   // at_close (split.cu:131-150)
        gpu(mgv_stream_synchronize(last_stream));
        opened = false;
        // after_close closes every parent (split.cu:152-163)
        for (d_operator_t *p : splitter->parents) p->close();
        return;
    }
    gpu(mgv_stream_synchronize(last_stream));
    gpu(mgv_device_synchronize());
    if (sink) {
        // at_close: the end of the rows stream, or the scalar result
        if (kind == SELECT_MAT) gpu(mgv_stream_out_close(sink->rows_out, nullptr));
        else gpu(mgv_ring_emit(sink->ring, 0, 0, res, nullptr));
        gpu(mgv_stream_synchronize(nullptr));
        // after_close: join the catcher, close the host parent
        if (sink->catcher.joinable()) sink->catcher.join();
        if (!sink->failure.empty()) throw mgv::error(MGV_E_INVALID, "gpu_to_cpu: " + sink->failure);
        if (sink->dropped)
            throw mgv::error(MGV_E_CAPACITY, "gpu_to_cpu: " + std::to_string(sink->dropped) +
                                                 " rows found no free device buffer in time");
        sink->parent->close();
    }
    opened = false;
}

d_operator_t::~d_operator_t() {
    for (auto &w : ws) mgv_free(w.second.first);
    mgv_free(res);
    mgv_free(stage);
    mgv_free(cursors);
}

void d_operator_t::destroy(d_operator_t *op) {
    if (!op) return;
    std::visit([](auto *p) { delete p; }, op->op);
    delete op;
}

// ------------------------------------------------------------------ h_operator_t

void h_operator_t::open() {
    std::visit([](auto *p) { p->open(); }, op);
}

void h_operator_t::consume(buffer_t *b) {
    std::visit([b](auto *p) { p->consume(b); }, op);
}

void h_operator_t::close() {
    std::visit([](auto *p) { p->close(); }, op);
}

void h_operator_t::destroy(h_operator_t *op) {
    if (!op) return;
    std::visit([](auto *p) { delete p; }, op->op);
    delete op;
}

// ------------------------------------------------------------------ generator

generator::generator(h_operator_t *parent_, int32_t *src, size_t N_)
    : parent(parent_), cols{src}, N(N_), vector_rows(buffer_t::capacity()), device(-1) {}

generator::generator(h_operator_t *parent_, std::vector<int32_t *> cols_, size_t N_, uint32_t vector_rows_, int dev)
    : parent(parent_), cols(std::move(cols_)), N(N_), vector_rows(vector_rows_), device(dev) {
    if (cols.empty() || cols.size() > (size_t)buffer_t::kMaxCols)
        throw mgv::error(MGV_E_INVALID, "generator: 1 to 4 columns");
    if (vector_rows == 0 || vector_rows % (4 * WARPSIZE) || vector_rows > buffer_t::capacity())
        throw mgv::error(MGV_E_INVALID, "generator: vector rows must be a multiple of 256 up to the buffer capacity");
}

void generator::open() { parent->open(); }

void generator::consume(buffer_t *) { throw mgv::error(MGV_E_INVALID, "generator::consume: a generator has no input"); }

void generator::close() {
    size_t off = 0;
    while (off < N) {
        buffer_t *b = buffer_manager<int32_t>::get_buffer();
        const uint32_t m = (uint32_t)std::min<size_t>(N - off, vector_rows);
        int32_t *c[buffer_t::kMaxCols];
        for (size_t i = 0; i < cols.size(); ++i) c[i] = cols[i] + off;
        b->set_columns(c, (int)cols.size());     // the vector aliases the columns (generators.cu:71-72)
        b->cnt = m;
        b->device = device;
        parent->consume(b);
        off += m;
    }
    parent->close();
}

// ------------------------------------------------------------------ materializer

materializer::materializer(int32_t *dst_) : dst(dst_) {}

void materializer::consume(buffer_t *data) {
    std::lock_guard<std::mutex> lk(mu);
    const uint32_t n = data->count();
    if (dst && n) {
        if (data->device >= 0) gpu(mgv_memcpy(dst + size, data->data, (size_t)n * sizeof(int32_t)));
        else memcpy(dst + size, data->data, (size_t)n * sizeof(int32_t));
    }
    size += n;
    if (data->scalar) {
        result64 += data->value64;
        results += 1;
    }
    // buffers of the manager go back to it (materializer.cu:34); the host
    // buffers a gpu_to_cpu hands over stay owned by that gpu_to_cpu
    if (buffer_manager<int32_t>::owns(data)) buffer_manager<int32_t>::release_buffer(data);
}

// ------------------------------------------------------------------ union_all_cpu / mem_move

union_all_cpu::union_all_cpu(h_operator_t *parent_, int children_) : parent(parent_), children(children_) {
    if (!parent) throw mgv::error(MGV_E_INVALID, "union_all_cpu: null parent");
    if (children < 1) throw mgv::error(MGV_E_INVALID, "union_all_cpu: at least one child");
}

void union_all_cpu::open() {
    std::lock_guard<std::mutex> lk(m);
    if (++children_opened == 1) parent->open();
}

void union_all_cpu::consume(buffer_t *data) { parent->consume(data); }

void union_all_cpu::close() {
    std::lock_guard<std::mutex> lk(m);
    if (++children_closed == children) {
        parent->close();
        children_closed = 0;   // ready for the next open / close round
        children_opened = 0;
    }
}

mem_move::mem_move(h_operator_t *parent_) : parent(parent_) {
    if (!parent) throw mgv::error(MGV_E_INVALID, "mem_move: null parent");
}

void mem_move::open() { parent->open(); }

void mem_move::consume(buffer_t *data) {
    if (data->device >= 0 && !data->scalar) {
        buffer_t *host = buffer_manager<int32_t>::get_buffer();
        const uint32_t n = data->count();
        if (n) gpu(mgv_memcpy(host->data, data->data, (size_t)n * sizeof(int32_t)));
        host->cnt = n;
        if (buffer_manager<int32_t>::owns(data)) buffer_manager<int32_t>::release_buffer(data);
        data = host;
    }
    parent->consume(data);
}

void mem_move::close() { parent->close(); }

// ------------------------------------------------------------------ consumer

consumer::consumer(p_operator_t parent_, const launch_conf &conf_) : parent(parent_), conf(conf_) {}

consumer::~consumer() {
    for (int i = 0; i < 2; ++i) {
        mgv_event_destroy(copied[i]);
        mgv_event_destroy(used[i]);
    }
    mgv_stream_destroy(strm);
    mgv_stream_destroy(copy_strm);
}

void consumer::open() {
    gpu(mgv_stream_create(&strm));
    gpu(mgv_stream_create(&copy_strm));
    for (int i = 0; i < 2; ++i) {
        gpu(mgv_event_create(&copied[i]));
        gpu(mgv_event_create(&used[i]));
        gpu(mgv_event_record(used[i], strm));
    }
    if (parent.d) parent.d->open();
    else parent.h->open();
}

void consumer::consume(buffer_t *b) {
    if (!parent.d) {   // host parent: pass through (exchange.cu:286-289)
        parent.h->consume(b);
        return;
    }
    if (b->device >= 0) {   // resident vector: launch on it directly
        parent.d->consume(b, strm);
        buffer_manager<int32_t>::release_buffer(b);
        ++nvec;
        return;
    }
    // host vector: copy into the next staging buffer (double buffered), then launch
    const int i = (int)(nvec & 1);
    if (!stage[i]) stage[i] = buffer_manager<int32_t>::h_get_buffer(0);
    const uint32_t n = b->count();
    const uint32_t pitch = (n + 255) / 256 * 256;
    if ((uint64_t)pitch * b->ncols > buffer_t::capacity())
        throw mgv::error(MGV_E_CAPACITY, "consumer: vector does not fit a device buffer");
    gpu(mgv_stream_wait_event(copy_strm, used[i]));
    int32_t *c[buffer_t::kMaxCols];
    for (int k = 0; k < b->ncols; ++k) {
        c[k] = stage[i]->cols[0] + (size_t)k * pitch;
        gpu(mgv_memcpy_async(c[k], b->cols[k], (size_t)n * sizeof(int32_t), copy_strm));
    }
    gpu(mgv_event_record(copied[i], copy_strm));
    buffer_t v = *stage[i];
    v.set_columns(c, b->ncols);
    v.cnt = n;
    v.device = 0;
    gpu(mgv_stream_wait_event(strm, copied[i]));
    parent.d->consume(&v, strm);
    gpu(mgv_event_record(used[i], strm));
    // the host vector may be reused once its copy is done
    gpu(mgv_event_synchronize(copied[i]));
    buffer_manager<int32_t>::release_buffer(b);
    ++nvec;
}

void consumer::close() {
    gpu(mgv_stream_synchronize(strm));
    if (parent.d) parent.d->close();
    else parent.h->close();
    for (int i = 0; i < 2; ++i) {
        if (stage[i]) buffer_manager<int32_t>::release_buffer(stage[i]);
        stage[i] = nullptr;
    }
}

// ------------------------------------------------------------------ exchange

exchange::exchange(const std::vector<p_operator_t> &parents, const std::vector<launch_conf> &parent_conf) {
    if (parents.size() != parent_conf.size()) throw mgv::error(MGV_E_INVALID, "exchange: parents/confs mismatch");
    for (size_t i = 0; i < parents.size(); ++i) consumers.emplace_back(new consumer(parents[i], parent_conf[i]));
    for (auto &c : consumers) firers.emplace_back(&exchange::fire, this, c.get());
}

exchange::~exchange() {
    {
        std::lock_guard<std::mutex> lk(mu);
        remaining_producers = 0;
    }
    cv.notify_all();
    for (auto &t : firers)
        if (t.joinable()) t.join();
    // vectors left by a failed run go back to the manager
    for (buffer_t *b : ready)
        if (buffer_manager<int32_t>::owns(b)) buffer_manager<int32_t>::release_buffer(b);
    ready.clear();
}

void exchange::fire(consumer *c) {
    try {
        c->open();
        for (;;) {
            buffer_t *b = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !ready.empty() || remaining_producers <= 0; });
                if (ready.empty()) break;
                b = ready.back();   // LIFO like the reference's ready_pool (exchange.cu:98)
                ready.pop_back();
            }
            c->consume(b);
        }
        c->close();
    } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!failure) failure = std::current_exception();
        remaining_producers = 0;
        // hand queued vectors back so a producer blocked on a free buffer
        // wakes up and sees the failure
        for (buffer_t *b : ready)
            if (buffer_manager<int32_t>::owns(b)) buffer_manager<int32_t>::release_buffer(b);
        ready.clear();
        cv.notify_all();
    }
}

void exchange::consume(buffer_t *b) {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (failure) {
            // a consumer failed: stop the producer instead of queueing
            // vectors nobody will release
            if (buffer_manager<int32_t>::owns(b)) buffer_manager<int32_t>::release_buffer(b);
            std::rethrow_exception(failure);
        }
        ready.push_back(b);
    }
    cv.notify_one();
}

void exchange::close() {
    {
        std::lock_guard<std::mutex> lk(mu);
        --remaining_producers;
    }
    cv.notify_all();
    for (auto &t : firers)
        if (t.joinable()) t.join();
    if (failure) std::rethrow_exception(failure);
}
