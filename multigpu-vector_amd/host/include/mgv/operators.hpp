// operators.hpp -- the operator API of the C++ drop-in layer.
//
// Same names, arguments and lifecycle as the reference's operator framework
// (src/operators/operator.cuh:42-124) and operators:
//   device side (one per-vector launch per operator chain):
//     unstable_select<WARPSIZE, less_eq_than<int32_t>, int32_t>   select3.cuh:10-49
//     aggregation<WARPSIZE, 0>                                    aggregation.cuh:8-30
//     hashjoin<WARPSIZE> / hashjoin_builder<WARPSIZE>             hashjoin.cuh:8-58
//     gpu_to_cpu<WARPSIZE, 64, buffer_t *>                        gpu_to_cpu.cuh:14-64
//     split<WARPSIZE, int32_t> / union_all<WARPSIZE, int32_t>     split.cuh:10-35, union_all.cuh:11-37
//   host side:
//     generator, exchange (+ consumer), materializer              generators.cuh, exchange.cuh, materializer.cuh
//     union_all_cpu, mem_move                                     union_all_cpu.cuh, mem_move.cuh
//
// d_operator_t::create<Op>(conf, args...) builds a host-side operator
// description; open() resolves the chain from an operator up to its
// gpu_to_cpu sink into one libmgv pipeline (select -> sum, select ->
// materialise, [select ->] probe -> count / sum, [select ->] build), so a
// vector goes through the whole chain in one kernel launch, as in
// launch_consume_pipeline (exchange.cu:124-194).  The reference dispatched a
// `variant` of device operator pointers per 128-row chunk; here the chain is
// fixed once per query and the device code is statically composed.
//
// Differences a caller sees:
//  - results are exact: SUM / COUNT are int64 (gpu_to_cpu::result(),
//    materializer::result64); the int32 value the reference would have
//    delivered (low 32 bits) is what the host buffers carry;
//  - vectors may carry up to 4 aligned columns (generator with several
//    columns): select tests column 0 and passes the remaining columns (or
//    column 0 when alone) to its parent; the probe joins on the first column
//    it receives and sums the next one plus the build payload;
//  - a split stages its parents' rows and hands them on as vectors when its
//    staging fills and at close (the reference forwarded every 128 rows),
//    partitioning on the first column it receives;
//  - the chains of a union_all's children share its accumulators (one result
//    slot, one row buffer), and each child opens / closes it once;
//  - mem_move copies the first column of a device vector to a pinned host
//    buffer (the vectors a gpu_to_cpu hands on are host vectors already).
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <variant>
#include <vector>

#include "buffer_manager.hpp"

typedef buffer_pool<int32_t>::buffer_t buffer_t;

class d_operator_t;
class h_operator_t;

// A parent of a host operator: device or host (operator.cuh:117-124 is a
// union; the tag replaces the reference's get_device(pointer) test).
struct p_operator_t {
    d_operator_t *d = nullptr;
    h_operator_t *h = nullptr;
    p_operator_t() = default;
    p_operator_t(d_operator_t *d_) : d(d_) {}
    p_operator_t(h_operator_t *h_) : h(h_) {}
};

template <typename T>
struct less_eq_than {
    T bound;
    constexpr less_eq_than(T b) : bound(b) {}
    constexpr bool operator()(const T &x) const { return x <= bound; }
};

// ------------------------------------------------------------ device operators

template <size_t warp_size = WARPSIZE, typename F = less_eq_than<int32_t>, typename... T>
class unstable_select {
public:
    d_operator_t *parent;
    F filt;
    unstable_select(d_operator_t *parent_, F f, int grid_size = 0, int dev = 0) : parent(parent_), filt(f) {
        (void)grid_size;
        (void)dev;
    }
};

template <size_t warp_size = WARPSIZE, int32_t neutral_value = 0>
class aggregation {
    static_assert(neutral_value == 0, "only SUM (neutral value 0) is implemented");

public:
    d_operator_t *parent;
    int32_t shared_offset;
    aggregation(d_operator_t *parent_, int32_t shared_offset_, int grid_size = 0, int dev = 0)
        : parent(parent_), shared_offset(shared_offset_) {
        (void)grid_size;
        (void)dev;
    }
};

// Open-addressing table shared by a hashjoin and its builder.
struct mgv_table_state {
    uint64_t *slots = nullptr;
    uint32_t nbuckets = 0;
    uint64_t max_size = 0;
    ~mgv_table_state() { mgv_free(slots); }
};

template <size_t warp_size = WARPSIZE>
class hashjoin_builder {
public:
    std::shared_ptr<mgv_table_state> table;
    explicit hashjoin_builder(std::shared_ptr<mgv_table_state> t) : table(std::move(t)) {}
};

template <size_t warp_size = WARPSIZE>
class hashjoin {
public:
    d_operator_t *parent;
    std::shared_ptr<mgv_table_state> table;
    d_operator_t *builder;
    // The table holds max_size keys at load factor 0.5 and has at least
    // 2^log_table_size slots (the reference's 2^log_table_size chain heads,
    // hashjoin.cu:63-71); log_table_size must be in [0, 31].
    hashjoin(d_operator_t *parent_, int32_t log_table_size, int32_t max_size, const launch_conf &conf);
    ~hashjoin();   // destroys the builder
};

// Hands a chain's results to its host parent (gpu_to_cpu.cu:7-184).  Rows
// (select -> gpu_to_cpu) stream through a small pool of this operator's own
// output buffers in pinned host memory: kernels append them to the current
// pool buffer, swap in a fresh one when it is full, and publish each full
// buffer on a ring; the catcher thread copies it into a pinned host buffer
// of the buffer manager, hands the pool buffer straight back (no kernel, no
// copy engine), and passes the manager's buffer to the parent, which owns it
// from then on.  A scalar result (aggregation / join count) arrives on the
// same ring.
template <size_t warp_size = WARPSIZE, size_t size = 64, typename T = buffer_t *>
class gpu_to_cpu {
public:
    h_operator_t *parent;
    int device;
    static constexpr uint32_t kOutBuffers = 8, kOutRows = 1u << 20;
    mgv_pool *out_pool = nullptr;           // kOutBuffers pinned host buffers of kOutRows rows
    mgv_ring *ring = nullptr;
    mgv_stream_out *rows_out = nullptr;
    std::thread catcher;
    std::atomic<bool> stop{false};          // ends a catcher whose chain failed
    mgv_result last{};                      // count = rows handed over, or the scalar result
    uint64_t buffers = 0, dropped = 0;      // of the last rows stream
    std::string failure;                    // set by the catcher
    std::vector<int32_t> rows;              // scalar result row
    std::vector<buffer_t> out;              // scalar buffer handed to the parent

    gpu_to_cpu(h_operator_t *parent_, int device_);
    ~gpu_to_cpu();
    const mgv_result &result() const { return last; }
};

// Routes each row to parent (x & 1) of its first column x (split.cu:49-118;
// mgv_split): rows whose destination has no parent are dropped, parents
// beyond the second receive nothing.  A parent may appear several times.
template <size_t warp_size = WARPSIZE, typename T = int32_t>
class split {
public:
    std::vector<d_operator_t *> parents;
    split(std::vector<d_operator_t *> parents_, launch_conf conf, int dev = 0);
};

// Fan-in of num_of_children children into one parent: the first child's
// open opens the parent and the last child's close closes it
// (union_all.cu:19-30, 63-74); children may consume from several threads.
template <size_t warp_size = WARPSIZE, typename T = int32_t>
class union_all {
public:
    d_operator_t *parent;
    int num_of_children;
    int num_of_active_children = 0;
    int num_of_closed_children = 0;
    std::mutex host_lock;
    union_all(d_operator_t *parent_, int num_of_children_, launch_conf conf, int dev = 0);
};

typedef unstable_select<WARPSIZE, less_eq_than<int32_t>, int32_t> select_t;
typedef aggregation<WARPSIZE, 0> aggregation_t;
typedef hashjoin<WARPSIZE> hashjoin_t;
typedef hashjoin_builder<WARPSIZE> hashjoin_builder_t;
typedef gpu_to_cpu<WARPSIZE, 64, buffer_t *> gpu_to_cpu_t;
typedef split<WARPSIZE, int32_t> split_t;
typedef union_all<WARPSIZE, int32_t> union_all_t;

class d_operator_t {
public:
    typedef std::variant<select_t *, aggregation_t *, hashjoin_builder_t *, hashjoin_t *, gpu_to_cpu_t *, split_t *,
                         union_all_t *>
        op_t;
    op_t op;
    launch_conf conf;

    template <typename Op, typename... Args>
    static d_operator_t *create(launch_conf lc, Args... args) {
        return new d_operator_t(lc, op_t(new Op(args...)));
    }
    static void destroy(d_operator_t *op);

    template <typename T>
    T get() {
        return std::get<T>(op);
    }

    // host lifecycle (operator.cu:126-135, 150-159)
    void open();
    void close();
    // one vector through the chain (launch_consume_pipeline, exchange.cu:160-194)
    void consume(buffer_t *vec, mgv_stream_t s);

    ~d_operator_t();

private:
    d_operator_t(launch_conf lc, op_t o) : op(std::move(o)), conf(lc) {}

    enum Kind { NONE, SELECT_SUM, SELECT_MAT, PROBE, SELECT_PROBE, BUILD, SELECT_BUILD, SPLIT };
    Kind kind = NONE;
    int32_t bound = 0;
    bool sum = false;
    bool filtered = false;               // a select heads the chain
    std::shared_ptr<mgv_table_state> table;
    gpu_to_cpu_t *sink = nullptr;
    mgv_result *res = nullptr;           // device
    mgv_stream_t last_stream = nullptr;
    bool opened = false;
    // a union_all further up the chain: this chain shares its result slot
    // and materialised rows, and opens / closes it as one of its children
    d_operator_t *through = nullptr;
    // split: per-destination staging columns and their append cursors
    split_t *splitter = nullptr;
    int32_t *stage = nullptr;            // device, [dest][col][stage_cap]
    uint64_t *cursors = nullptr;         // device, [dest]
    uint64_t stage_cap = 0, stage_upper = 0;
    int stage_cols = 0;
    // join vectors large enough for the partitioned LDS paths: their
    // workspace, one per stream the chain is fed on
    static constexpr uint64_t kLdsRows = 1ull << 20, kLdsBuildRows = 1ull << 21;
    static constexpr uint64_t kLdsTableBytes = 64ull << 20;
    static constexpr uint32_t kLdsFanout = 32;
    std::mutex ws_mu;
    std::map<mgv_stream_t, std::pair<void *, uint64_t>> ws;

    void resolve();
    void do_open();
    void do_close();
    void flush_split(mgv_stream_t s);
    uint32_t split_dests() const;
    void *workspace(mgv_stream_t s, uint64_t bytes);
    bool lds_join(uint64_t n, bool build, uint64_t *need) const;
};

// ------------------------------------------------------------ host operators

class generator;
class materializer;
class exchange;
class union_all_cpu;
class mem_move;

class h_operator_t {
public:
    typedef std::variant<generator *, materializer *, exchange *, union_all_cpu *, mem_move *> op_t;
    op_t op;

    template <typename Op, typename... Args>
    static h_operator_t *create(Args... args) {
        return new h_operator_t(op_t(new Op(args...)));
    }
    static void destroy(h_operator_t *op);

    void open();
    void consume(buffer_t *buff);
    void close();

    template <typename T>
    T get() {
        return std::get<T>(op);
    }

private:
    explicit h_operator_t(op_t o) : op(std::move(o)) {}
};

// Slices columns into vectors of `vector_rows` rows (default: the buffer
// capacity) and pushes them to its parent (generators.cu:55-90).  The
// columns may be device or host memory; a vector aliases its column memory.
class generator {
public:
    h_operator_t *parent;
    std::vector<int32_t *> cols;
    size_t N;
    uint32_t vector_rows;
    int device;     // -1: host columns

    generator(h_operator_t *parent, int32_t *src, size_t N);
    generator(h_operator_t *parent, std::vector<int32_t *> cols, size_t N, uint32_t vector_rows, int device);
    void open();
    void consume(buffer_t *);
    void close();
};

// Host sink: keeps the data of every buffer it receives (main9.cu:633 reads
// dst/size) and releases the buffer (materializer.cu:10-40).
class materializer {
public:
    int32_t *dst;
    size_t size = 0;
    int64_t result64 = 0;       // exact value of scalar results
    uint64_t results = 0;       // scalar results received
    std::mutex mu;

    explicit materializer(int32_t *dst);
    void open() {}
    void consume(buffer_t *data);
    void close() {}
};

// Host fan-in of `children` producers into one parent: the first child's
// open opens the parent, the last child's close closes it; vectors go
// straight through (union_all_cpu.cu:3-19).  The parent must accept
// consumes from several threads (materializer and exchange do).
class union_all_cpu {
public:
    h_operator_t *parent;
    const int children;
    int children_closed = 0;
    int children_opened = 0;
    std::mutex m;

    union_all_cpu(h_operator_t *parent, int children);
    void open();
    void consume(buffer_t *data);
    void close();
};

// Hands its parent host vectors: a device vector's rows (first column) are
// copied into a pinned host buffer of the buffer manager, and the device
// buffer goes back to the manager when it owns it (mem_move.cu:10-50).
class mem_move {
public:
    h_operator_t *parent;

    explicit mem_move(h_operator_t *parent);
    void open();
    void consume(buffer_t *data);
    void close();
};

// Per-GPU consumer of an exchange: routes vectors to a device operator on
// its own streams, copying host vectors to the device first with double
// buffering (exchange.cu:196-318).
class consumer {
public:
    p_operator_t parent;
    launch_conf conf;
    mgv_stream_t strm = nullptr, copy_strm = nullptr;
    mgv_event_t copied[2] = {nullptr, nullptr};
    mgv_event_t used[2] = {nullptr, nullptr};
    buffer_t *stage[2] = {nullptr, nullptr};
    uint64_t nvec = 0;

    consumer(p_operator_t parent, const launch_conf &conf);
    ~consumer();
    void open();
    void consume(buffer_t *buff);
    void close();
};

// Vector router (exchange.cu:9-109): a ready queue and one firer thread per
// consumer; close() ends the stream and joins the firers.
class exchange {
public:
    exchange(const std::vector<p_operator_t> &parents, const std::vector<launch_conf> &parent_conf);
    ~exchange();
    void open() {}
    void consume(buffer_t *buff);
    void close();

private:
    void fire(consumer *c);
    std::deque<buffer_t *> ready;
    std::mutex mu;
    std::condition_variable cv;
    int remaining_producers = 1;
    std::vector<std::unique_ptr<consumer>> consumers;
    std::vector<std::thread> firers;
    std::exception_ptr failure;
};
