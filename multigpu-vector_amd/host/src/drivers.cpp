// drivers.cpp -- extern "C" entry points of libmgv_ops.so: the reference's
// plan builders (src/main9.cu:150-367) assembled from the C++ operators and
// run with generator open/close (src/main9.cu:604-614), plus the plain
// per-vector launch path of a query batch.
#include <climits>
#include <string>

#include "../../../include/mgv_ops.h"
#include "mgv/operators.hpp"

namespace mgvhost {
void serve_run(const mgv_serve_request *requests, uint32_t n, const mgv_serve_config *config,
               mgv_serve_result *results, mgv_serve_stats *stats);
}

namespace {

thread_local std::string g_err;

template <class F>
int guarded(const char *what, F &&f) {
    try {
        f();
        return MGV_OK;
    } catch (const mgv::error &e) {
        g_err = std::string(what) + ": " + e.what();
        return e.code ? e.code : MGV_E_INVALID;
    } catch (const std::exception &e) {
        g_err = std::string(what) + ": " + e.what();
        return MGV_E_INVALID;
    }
}

launch_conf make_conf(uint32_t grid, uint32_t block) {
    if (grid == 0 || (block != 64 && block != 256 && block != 1024))
        throw mgv::error(MGV_E_INVALID, "grid must be > 0 and block one of 64, 256, 1024");
    return launch_conf(dim3(grid), dim3(block), 0, 0);
}

void check_vector_rows(uint32_t vector_rows) {
    if (vector_rows == 0 || vector_rows % 256 || vector_rows > buffer_t::capacity())
        throw mgv::error(MGV_E_INVALID, "vector_rows must be a positive multiple of 256 up to 4 Mi");
}

void require_manager() {
    if (!buffer_manager<int32_t>::initialized()) throw mgv::error(MGV_E_NOT_INIT, "mgv_ops_init was not called");
}

// The sink side every plan shares: gpu_to_cpu -> exchange -> materializer.
struct Sink {
    h_operator_t *mat = nullptr, *xroot = nullptr;
    explicit Sink(int32_t *dst) {
        const launch_conf mat_conf(dim3(1), dim3(1), 0, -1);
        mat = h_operator_t::create<materializer>(dst);
        xroot = h_operator_t::create<exchange>(std::vector<p_operator_t>{p_operator_t(mat)},
                                               std::vector<launch_conf>{mat_conf});
    }
    ~Sink() {
        h_operator_t::destroy(xroot);
        h_operator_t::destroy(mat);
    }
    materializer *m() { return mat->get<materializer *>(); }
};

// generator -> exchange -> heads (one consumer, with its own streams, per
// head); run with open / close.
void run_generator(const std::vector<d_operator_t *> &heads, const launch_conf &conf, std::vector<int32_t *> cols,
                   uint64_t n, uint32_t vector_rows, int device) {
    std::vector<p_operator_t> parents;
    for (d_operator_t *h : heads) parents.emplace_back(h);
    h_operator_t *xgen = h_operator_t::create<exchange>(parents, std::vector<launch_conf>(heads.size(), conf));
    try {
        generator gen(xgen, std::move(cols), n, vector_rows, device);
        gen.open();
        gen.close();
    } catch (...) {
        h_operator_t::destroy(xgen);
        throw;
    }
    h_operator_t::destroy(xgen);
}

void run_generator(d_operator_t *head, const launch_conf &conf, std::vector<int32_t *> cols, uint64_t n,
                   uint32_t vector_rows, int device) {
    run_generator(std::vector<d_operator_t *>{head}, conf, std::move(cols), n, vector_rows, device);
}

// generator -> exchange -> {select(thres)} x nstreams -> union_all(nstreams)
// -> [aggregation ->] gpu_to_cpu -> exchange -> materializer.
void run_fan_in(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                int device, const launch_conf &conf, uint32_t nstreams, bool aggregate, Sink &sink,
                mgv_result *result) {
    if (nstreams == 0 || nstreams > 16) throw mgv::error(MGV_E_INVALID, "nstreams must be in [1, 16]");
    d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, sink.xroot, conf.device);
    d_operator_t *aggr =
        aggregate ? d_operator_t::create<aggregation_t>(conf, g2c, 8000, conf.get_blocks_per_grid(), 0) : nullptr;
    d_operator_t *u = d_operator_t::create<union_all_t>(conf, aggregate ? aggr : g2c, (int)nstreams, conf,
                                                        conf.device);
    std::vector<d_operator_t *> sels;
    auto cleanup = [&] {
        for (d_operator_t *x : sels) d_operator_t::destroy(x);
        d_operator_t::destroy(u);
        d_operator_t::destroy(aggr);
        d_operator_t::destroy(g2c);
    };
    try {
        for (uint32_t i = 0; i < nstreams; ++i)
            sels.push_back(d_operator_t::create<select_t>(conf, u, less_eq_than<int32_t>(thres),
                                                          conf.get_blocks_per_grid(), 0));
        run_generator(sels, conf, {(int32_t *)pred, (int32_t *)val}, n, vector_rows, device);
    } catch (...) {
        cleanup();
        throw;
    }
    *result = g2c->get<gpu_to_cpu_t *>()->result();
    cleanup();
}

}  // namespace

extern "C" {

const char *mgv_ops_last_error(void) { return g_err.c_str(); }

int mgv_ops_init(uint32_t nbuffers, uint32_t buff_buffer_size, uint32_t keep_threshold) {
    return guarded("mgv_ops_init", [&] {
        buffer_manager<int32_t>::init((int)nbuffers, (int)buff_buffer_size, (int)keep_threshold);
    });
}

int mgv_ops_shutdown(void) {
    return guarded("mgv_ops_shutdown", [&] { buffer_manager<int32_t>::destroy(); });
}

int mgv_ops_pool_stats(uint32_t *device_free, uint32_t *device_cached, uint32_t *device_total, uint32_t *host_free) {
    return guarded("mgv_ops_pool_stats", [&] {
        require_manager();
        uint32_t avail = 0, prod = 0, stacked = 0;
        buffer_manager<int32_t>::device_pool()->stats(&avail, &prod, &stacked);
        if (device_free) *device_free = stacked;
        if (device_cached) *device_cached = buffer_manager<int32_t>::device_cached();
        if (device_total) *device_total = (uint32_t)buffer_manager<int32_t>::device_pool()->size();
        if (host_free) *host_free = buffer_manager<int32_t>::host_available();
    });
}

int mgv_ops_sum_select(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                       int device, uint32_t grid, uint32_t block, int32_t *dst, int64_t *sum) {
    return guarded("mgv_ops_sum_select", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !sum) throw mgv::error(MGV_E_INVALID, "null argument");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        // create_sum_select_pipeline (main9.cu:214-217)
        d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, sink.xroot, conf.device);
        d_operator_t *aggr = d_operator_t::create<aggregation_t>(conf, g2c, 8000, conf.get_blocks_per_grid(), 0);
        d_operator_t *sel = d_operator_t::create<select_t>(conf, aggr, less_eq_than<int32_t>(thres),
                                                           conf.get_blocks_per_grid(), 0);
        auto cleanup = [&] {
            d_operator_t::destroy(sel);
            d_operator_t::destroy(aggr);
            d_operator_t::destroy(g2c);
        };
        try {
            run_generator(sel, conf, {(int32_t *)pred, (int32_t *)val}, n, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        *sum = n ? sink.m()->result64 : 0;
        if (dst && n == 0) dst[0] = 0;
        cleanup();
    });
}

int mgv_ops_select(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, int apply_select,
                   uint32_t vector_rows, int device, uint32_t grid, uint32_t block, int32_t *dst, uint64_t dst_cap,
                   uint64_t *count) {
    return guarded("mgv_ops_select", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !count) throw mgv::error(MGV_E_INVALID, "null argument");
        if (dst && dst_cap < n) throw mgv::error(MGV_E_CAPACITY, "dst must hold n rows");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, sink.xroot, conf.device);
        d_operator_t *sel = apply_select ? d_operator_t::create<select_t>(conf, g2c, less_eq_than<int32_t>(thres),
                                                                          conf.get_blocks_per_grid(), 0)
                                         : nullptr;
        auto cleanup = [&] {
            d_operator_t::destroy(sel);
            d_operator_t::destroy(g2c);
        };
        try {
            std::vector<int32_t *> cols;
            if (apply_select) cols = {(int32_t *)pred, (int32_t *)val};
            else cols = {(int32_t *)val};
            run_generator(sel ? sel : g2c, conf, cols, n, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        *count = n ? g2c->get<gpu_to_cpu_t *>()->result().count : 0;
        cleanup();
    });
}

int mgv_ops_select_split(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                         int device, uint32_t grid, uint32_t block, int32_t *dst, uint64_t dst_cap, uint64_t *count) {
    return guarded("mgv_ops_select_split", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !count) throw mgv::error(MGV_E_INVALID, "null argument");
        if (dst && dst_cap < n) throw mgv::error(MGV_E_CAPACITY, "dst must hold n rows");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        // create_select_pipeline_exp (main9.cu:239-249): both parents of the
        // split are the same union_all
        d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, sink.xroot, conf.device);
        d_operator_t *g2cu = d_operator_t::create<union_all_t>(conf, g2c, 2, conf, conf.device);
        d_operator_t *spl = d_operator_t::create<split_t>(conf, std::vector<d_operator_t *>{g2cu, g2cu}, conf,
                                                          conf.device);
        d_operator_t *sel = d_operator_t::create<select_t>(conf, spl, less_eq_than<int32_t>(thres),
                                                           conf.get_blocks_per_grid(), 0);
        auto cleanup = [&] {
            d_operator_t::destroy(sel);
            d_operator_t::destroy(spl);
            d_operator_t::destroy(g2cu);
            d_operator_t::destroy(g2c);
        };
        try {
            run_generator(sel, conf, {(int32_t *)pred, (int32_t *)val}, n, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        *count = n ? g2c->get<gpu_to_cpu_t *>()->result().count : 0;
        cleanup();
    });
}

int mgv_ops_split_sums(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                       int device, uint32_t grid, uint32_t block, int64_t *sums, uint64_t *counts) {
    return guarded("mgv_ops_split_sums", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !sums || !counts) throw mgv::error(MGV_E_INVALID, "null argument");
        const launch_conf conf = make_conf(grid, block);
        Sink sink0(nullptr), sink1(nullptr);
        d_operator_t *g2c[2], *aggr[2];
        g2c[0] = d_operator_t::create<gpu_to_cpu_t>(conf, sink0.xroot, conf.device);
        g2c[1] = d_operator_t::create<gpu_to_cpu_t>(conf, sink1.xroot, conf.device);
        for (int p = 0; p < 2; ++p)
            aggr[p] = d_operator_t::create<aggregation_t>(conf, g2c[p], 8000, conf.get_blocks_per_grid(), 0);
        d_operator_t *spl = d_operator_t::create<split_t>(conf, std::vector<d_operator_t *>{aggr[0], aggr[1]}, conf,
                                                          conf.device);
        d_operator_t *sel = d_operator_t::create<select_t>(conf, spl, less_eq_than<int32_t>(thres),
                                                           conf.get_blocks_per_grid(), 0);
        auto cleanup = [&] {
            d_operator_t::destroy(sel);
            d_operator_t::destroy(spl);
            for (int p = 0; p < 2; ++p) {
                d_operator_t::destroy(aggr[p]);
                d_operator_t::destroy(g2c[p]);
            }
        };
        try {
            run_generator(sel, conf, {(int32_t *)pred, (int32_t *)val}, n, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        for (int p = 0; p < 2; ++p) {
            const mgv_result &r = g2c[p]->get<gpu_to_cpu_t *>()->result();
            sums[p] = n ? r.sum : 0;
            counts[p] = n ? r.count : 0;
        }
        cleanup();
    });
}

int mgv_ops_sum_select_streams(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres,
                               uint32_t vector_rows, int device, uint32_t grid, uint32_t block, uint32_t nstreams,
                               int64_t *sum) {
    return guarded("mgv_ops_sum_select_streams", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !sum) throw mgv::error(MGV_E_INVALID, "null argument");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(nullptr);
        mgv_result r{};
        run_fan_in(pred, val, n, thres, vector_rows, device, conf, nstreams, true, sink, &r);
        *sum = n ? sink.m()->result64 : 0;
    });
}

int mgv_ops_select_streams(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                           int device, uint32_t grid, uint32_t block, uint32_t nstreams, int32_t *dst,
                           uint64_t dst_cap, uint64_t *count) {
    return guarded("mgv_ops_select_streams", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !count) throw mgv::error(MGV_E_INVALID, "null argument");
        if (dst && dst_cap < n) throw mgv::error(MGV_E_CAPACITY, "dst must hold n rows");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        mgv_result r{};
        run_fan_in(pred, val, n, thres, vector_rows, device, conf, nstreams, false, sink, &r);
        *count = n ? r.count : 0;
    });
}

int mgv_ops_select_pipelines(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres,
                             uint32_t vector_rows, int device, uint32_t grid, uint32_t block, uint32_t npipelines,
                             int aggregate, int32_t *dst, uint64_t dst_cap, int64_t *value) {
    return guarded("mgv_ops_select_pipelines", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!pred || !val || !value) throw mgv::error(MGV_E_INVALID, "null argument");
        if (npipelines == 0 || npipelines > 16) throw mgv::error(MGV_E_INVALID, "npipelines must be in [1, 16]");
        if (dst && !aggregate && dst_cap < n) throw mgv::error(MGV_E_CAPACITY, "dst must hold n rows");
        if (dst && aggregate && dst_cap < npipelines)
            throw mgv::error(MGV_E_CAPACITY, "dst must hold one partial sum per pipeline");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        // create_select_pipeline / create_sum_select_pipeline with
        // npipelines launch confs (main9.cu:150-174, 202-226): one chain per
        // conf, their host outputs joined by union_all_cpu
        h_operator_t *xroot_pre =
            npipelines > 1 ? h_operator_t::create<union_all_cpu>(sink.xroot, (int)npipelines) : sink.xroot;
        std::vector<h_operator_t *> mmvs;
        std::vector<d_operator_t *> ops, heads;
        auto cleanup = [&] {
            for (auto it = ops.rbegin(); it != ops.rend(); ++it) d_operator_t::destroy(*it);
            for (h_operator_t *m : mmvs) h_operator_t::destroy(m);
            if (xroot_pre != sink.xroot) h_operator_t::destroy(xroot_pre);
        };
        try {
            for (uint32_t i = 0; i < npipelines; ++i) {
                h_operator_t *mmv = h_operator_t::create<mem_move>(xroot_pre);
                mmvs.push_back(mmv);
                d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, mmv, conf.device);
                ops.push_back(g2c);
                d_operator_t *up = g2c;
                if (aggregate) {
                    up = d_operator_t::create<aggregation_t>(conf, g2c, 8000, conf.get_blocks_per_grid(), 0);
                    ops.push_back(up);
                }
                d_operator_t *sel = d_operator_t::create<select_t>(conf, up, less_eq_than<int32_t>(thres),
                                                                   conf.get_blocks_per_grid(), 0);
                ops.push_back(sel);
                heads.push_back(sel);
            }
            run_generator(heads, conf, {(int32_t *)pred, (int32_t *)val}, n, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        *value = aggregate ? (n ? sink.m()->result64 : 0) : (int64_t)sink.m()->size;
        cleanup();
    });
}

int mgv_ops_count_join(const int32_t *R, const int32_t *pred_r, uint64_t rn, const int32_t *L, const int32_t *pred_l,
                       uint64_t ln, int32_t bound, uint32_t vector_rows, int device, uint32_t grid, uint32_t block,
                       int32_t *dst, uint64_t *count) {
    return guarded("mgv_ops_count_join", [&] {
        require_manager();
        check_vector_rows(vector_rows);
        if (!R || !L || !count) throw mgv::error(MGV_E_INVALID, "null argument");
        if (rn > (uint64_t)INT32_MAX) throw mgv::error(MGV_E_CAPACITY, "build side above 2^31-1 rows");
        const launch_conf conf = make_conf(grid, block);
        Sink sink(dst);
        // create_count_join_1G_pipeline (main9.cu:337-365); the select
        // variants put unstable_select in front of the builder / the probe.
        d_operator_t *g2c = d_operator_t::create<gpu_to_cpu_t>(conf, sink.xroot, conf.device);
        d_operator_t *hj = d_operator_t::create<hashjoin_t>(conf, g2c, 0, (int32_t)rn, conf);
        d_operator_t *builder = hj->get<hashjoin_t *>()->builder;
        d_operator_t *selb = pred_r ? d_operator_t::create<select_t>(conf, builder, less_eq_than<int32_t>(bound),
                                                                     conf.get_blocks_per_grid(), 0)
                                    : nullptr;
        d_operator_t *selp = pred_l ? d_operator_t::create<select_t>(conf, hj, less_eq_than<int32_t>(bound),
                                                                     conf.get_blocks_per_grid(), 0)
                                    : nullptr;
        auto cleanup = [&] {
            d_operator_t::destroy(selp);
            d_operator_t::destroy(selb);
            d_operator_t::destroy(hj);
            d_operator_t::destroy(g2c);
        };
        try {
            std::vector<int32_t *> bcols, pcols;
            if (pred_r) bcols = {(int32_t *)pred_r, (int32_t *)R};
            else bcols = {(int32_t *)R};
            if (pred_l) pcols = {(int32_t *)pred_l, (int32_t *)L};
            else pcols = {(int32_t *)L};
            run_generator(selb ? selb : builder, conf, bcols, rn, vector_rows, device);
            run_generator(selp ? selp : hj, conf, pcols, ln, vector_rows, device);
        } catch (...) {
            cleanup();
            throw;
        }
        *count = ln ? g2c->get<gpu_to_cpu_t *>()->result().count : 0;
        if (dst && ln == 0) dst[0] = 0;
        cleanup();
    });
}

int mgv_ops_run_plain(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, mgv_launch L,
                      mgv_stream_t s) {
    return guarded("mgv_ops_run_plain", [&] {
        check_vector_rows(vector_rows);
        if (!queries && nqueries) throw mgv::error(MGV_E_INVALID, "null queries");
        for (uint32_t qi = 0; qi < nqueries; ++qi) {
            const mgv_query &q = queries[qi];
            const int32_t bound = q.pred ? q.bound : INT32_MAX;
            for (uint64_t lo = 0; lo < q.n; lo += vector_rows) {
                const uint64_t m = std::min<uint64_t>(vector_rows, q.n - lo);
                const int32_t *pred = q.pred ? q.pred + lo : nullptr;
                const int32_t *key = q.key ? q.key + lo : nullptr;
                const int32_t *val = q.val ? q.val + lo : nullptr;
                switch (q.kind) {
                    case MGV_Q_SELECT_SUM:
                        gpu(mgv_select_sum(pred ? pred : val, val, m, bound, q.res, L, s));
                        break;
                    case MGV_Q_PROBE_SUM:
                        gpu(mgv_hashjoin_probe(q.slots, q.nbuckets, pred, q.bound, key, val, m, q.row_base + (int64_t)lo,
                                               q.flags, q.res, L, s));
                        break;
                    case MGV_Q_PROBE_GROUPBY:
                        gpu(mgv_hashjoin_probe_groupby(q.slots, q.nbuckets, pred, q.bound, key, val, m, q.gcnt, q.gsum,
                                                       q.ngroups, L, s));
                        break;
                    case MGV_Q_GROUPBY:
                        gpu(mgv_groupby(key, val, m, q.gcnt, q.gsum, q.ngroups, q.res, L, s));
                        break;
                    case MGV_Q_BUILD:
                        gpu(mgv_hashjoin_build(q.slots, q.nbuckets, key, val, q.payload_mode, pred, q.bound, m,
                                               q.row_base + (int64_t)lo, q.res, L, s));
                        break;
                    default:
                        throw mgv::error(MGV_E_INVALID, "unknown query kind");
                }
            }
        }
    });
}

int mgv_serve_run(const mgv_serve_request *requests, uint32_t n, const mgv_serve_config *config,
                  mgv_serve_result *results, mgv_serve_stats *stats) {
    return guarded("mgv_serve_run", [&] { mgvhost::serve_run(requests, n, config, results, stats); });
}

}  // extern "C"
