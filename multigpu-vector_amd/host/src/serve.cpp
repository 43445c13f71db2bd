// serve.cpp -- continuous-batching request loop over resident columns.
//
// One host thread drives the loop: it releases requests at their arrival
// time, admits waiting requests FIFO while their HBM footprint fits the free
// budget, packs the admitted requests into one batch per free stream slot
// (fused persistent launches, as mgv_batch_run), and takes completed
// requests off a device-written completion ring (each batch ends with a
// kernel that publishes every request's {id, result slot, result}),
// retiring a batch and returning its memory once all its requests are off
// the ring.  The reference has no
// serving layer; the per-query plan it would run is the select -> hashjoin ->
// aggregation chain of src/main9.cu:260-367 and the result hand-off of
// gpu_to_cpu (src/operators/gpu_to_cpu.cu:42-58,149-165).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/mgv_ops.h"
#include "mgv/common.hpp"

namespace {

using clock_t_ = std::chrono::steady_clock;

uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// First-fit allocator over [0, size) with coalescing free ranges.
class Arena {
public:
    explicit Arena(uint64_t size) : size_(size) { free_[0] = size; }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = align_up(std::max<uint64_t>(bytes, 1), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(std::max<uint64_t>(bytes, 1), 256);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t peak() const { return peak_; }
    uint64_t size() const { return size_; }

private:
    uint64_t size_, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

struct Admitted {
    uint32_t req;
    uint64_t off, bytes;     // arena reservation: table, or group arrays [+ group-by workspace]
    uint32_t nbuckets;
    uint64_t group_stage;    // offset (in groups) into the slot's pinned group staging
    uint64_t ws;             // workspace bytes at the end of the reservation: the partitioned
                             // group-by's, or a large join build's (LDS build); 0: fused launch
};

// Joins whose build side has at least this many rows build their table
// through LDS slices (mgv_hashjoin_build_lds, about twice as fast as the
// fused launch's CAS build for tables far larger than the L2) instead of in
// the batch's fused build launch.
constexpr uint64_t kLdsBuildRows = 1ull << 21;
constexpr uint32_t kLdsBuildFanout = 32;

// Work items of a request: a wave runs an item's rows one 256-row chunk
// after another (a few microseconds per chunk of random probes or
// compare-and-swaps), and every item ends with a flush whose atomics on the
// request's one result serialise (about 20 ns per item, measured).  A
// request of n rows in k items then takes about (n / k) * c + k * f with
// c / f ~ 1/2 per row, least at k = sqrt(n / 2): a 2^16-row build gets 128
// items, a 2^24-row probe about 2900, and from 2^27 rows on the items are
// capped at kMaxItems (every resident wave busy) or at vector_rows rows.
constexpr uint64_t kMaxItems = 8192;

// A batch's results are published together when its last launch ends, so a
// small request batched with a large one waits for the large one.  A batch
// takes requests while their rows (probe / group-by rows + build rows) add up
// to at most kBatchRows; a larger request runs in a batch of its own.
constexpr uint64_t kBatchRows = 1ull << 26;

struct Slot {
    mgv_stream_t stream = nullptr;
    bool busy = false;
    uint32_t pending = 0;             // requests of the batch not yet off the ring
    uint32_t batch = 0;
    std::vector<Admitted> q;
    std::vector<mgv_result> ring_res;   // the batch's results as they came off the ring
    uint32_t nbuild = 0, nmain = 0;
    uint64_t nbitems = 0, nmitems = 0;
    // device: [build queries | main queries | request ids | build items | main items], results, counters
    uint8_t *d_desc = nullptr;
    mgv_result *d_res = nullptr;
    uint32_t *d_counter = nullptr;
    // pinned host mirrors
    uint8_t *h_desc = nullptr;
    mgv_result *h_res = nullptr;
    uint64_t *h_groups = nullptr;     // per group-by query: cnt[ngroups] then sum[ngroups]
    uint64_t group_cap = 0;           // in groups
};

struct Server {
    const mgv_serve_request *R;
    uint32_t n;
    mgv_serve_config cfg;
    mgv_serve_result *out;
    Arena arena;
    uint8_t *d_arena = nullptr;
    std::vector<Slot> slots;
    std::vector<uint32_t> slot_of;    // request -> slot running it
    mgv_ring *ring = nullptr;         // completion ring, a slot per request that can be in flight
    uint64_t ring_results = 0;
    uint64_t items_cap = 0;
    size_t desc_bytes = 0;
    size_t ids_at = 0, items_at = 0;  // offsets in the descriptor staging

    Server(const mgv_serve_request *r, uint32_t n_, const mgv_serve_config &c, mgv_serve_result *o)
        : R(r), n(n_), cfg(c), out(o), arena(c.budget_bytes) {}

    ~Server() {
        mgv_device_synchronize();
        for (Slot &s : slots) {
            mgv_free(s.d_desc);
            mgv_free(s.d_res);
            mgv_free(s.d_counter);
            mgv_host_free(s.h_desc);
            mgv_host_free(s.h_res);
            mgv_host_free(s.h_groups);
            mgv_stream_destroy(s.stream);
        }
        mgv_ring_destroy(ring);
        mgv_free(d_arena);
    }

    // HBM bytes a request holds while it runs: its hash table, or its group
    // arrays plus, for many groups, the partitioned group-by's workspace.
    uint64_t footprint(const mgv_serve_request &q, uint32_t *nbuckets, uint64_t *ws = nullptr) const {
        if (q.kind == MGV_SERVE_JOIN_SUM) {
            uint64_t bytes = 0, w = 0;
            uint32_t nb = 0;
            gpu(mgv_ht_plan(q.bn, cfg.load_factor, &nb, &bytes));
            if (nbuckets) *nbuckets = nb;
            // a table whose buckets do not split into LDS slices builds in the fused launch
            if (q.bn < kLdsBuildRows || mgv_build_lds_workspace_size(q.bn, nb, kLdsBuildFanout, &w) != MGV_OK) w = 0;
            if (ws) *ws = w;
            return w ? align_up(bytes, 256) + w : bytes;
        }
        uint64_t w = 0;
        gpu(mgv_groupby_workspace_size(q.n, q.ngroups, &w));
        if (ws) *ws = w;
        return align_up((uint64_t)q.ngroups * 16, 256) + w;
    }

    // rows per work item of a request of `rows` rows: a multiple of 256 in
    // [256, vector_rows]
    uint32_t item_rows(uint64_t rows) const {
        const uint64_t k = std::max<uint64_t>(1, std::min<uint64_t>(kMaxItems, (uint64_t)std::sqrt(rows / 2.0)));
        const uint64_t want = align_up((rows + k - 1) / k, 256);
        return (uint32_t)std::min<uint64_t>(cfg.vector_rows, std::max<uint64_t>(256, want));
    }

    uint64_t items_of(uint64_t rows) const {
        const uint64_t v = item_rows(rows);
        return (rows + v - 1) / v;
    }

    // mgv_batch_plan per query with the query's own item size; the items
    // name their query's index in qs
    void plan(const mgv_query *qs, uint32_t nq, mgv_work_item *items, uint64_t cap, uint64_t *nitems) const {
        uint64_t k = 0;
        for (uint32_t i = 0; i < nq; ++i) {
            uint64_t c = 0;
            gpu(mgv_batch_plan(qs + i, 1, item_rows(qs[i].n), items + k, cap - k, &c));
            for (uint64_t j = 0; j < c; ++j) items[k + j].query = i;
            k += c;
        }
        *nitems = k;
    }

    void setup() {
        if (cfg.max_inflight < 1 || cfg.max_inflight > 8) throw mgv::error(MGV_E_INVALID, "max_inflight in 1..8");
        if (cfg.max_batch < 1 || cfg.max_batch > 1024) throw mgv::error(MGV_E_INVALID, "max_batch in 1..1024");
        if (cfg.vector_rows == 0 || cfg.vector_rows % 256) throw mgv::error(MGV_E_INVALID, "vector_rows % 256");
        std::vector<uint64_t> per;
        uint32_t max_groups = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const mgv_serve_request &q = R[i];
            if (q.kind != MGV_SERVE_JOIN_SUM && q.kind != MGV_SERVE_GROUPBY)
                throw mgv::error(MGV_E_INVALID, "request " + std::to_string(i) + ": unknown kind");
            if (!q.key || !q.val || (q.kind == MGV_SERVE_JOIN_SUM && (!q.bkey || !q.bval)) ||
                (q.kind == MGV_SERVE_GROUPBY && q.ngroups == 0))
                throw mgv::error(MGV_E_INVALID, "request " + std::to_string(i) + ": missing column");
            if (i && q.arrival_s < R[i - 1].arrival_s)
                throw mgv::error(MGV_E_INVALID, "requests must be sorted by arrival time");
            if (align_up(footprint(q, nullptr), 256) > cfg.budget_bytes)
                throw mgv::error(MGV_E_CAPACITY, "request " + std::to_string(i) + " exceeds the HBM budget");
            per.push_back(items_of(q.n) + (q.kind == MGV_SERVE_JOIN_SUM ? items_of(q.bn) : 0));
            if (q.kind == MGV_SERVE_GROUPBY) max_groups = std::max(max_groups, q.ngroups);
        }
        std::sort(per.rbegin(), per.rend());
        for (uint32_t i = 0; i < per.size() && i < cfg.max_batch; ++i) items_cap += per[i];
        items_cap = std::max<uint64_t>(items_cap, 1);
        ids_at = align_up(2 * (size_t)cfg.max_batch * sizeof(mgv_query), 256);
        items_at = align_up(ids_at + (size_t)cfg.max_batch * sizeof(uint64_t), 256);
        desc_bytes = items_at + (size_t)items_cap * sizeof(mgv_work_item);
        slot_of.assign(n, 0);
        gpu(mgv_ring_create(cfg.max_inflight * cfg.max_batch + 2, &ring));
        void *p = nullptr;
        gpu(mgv_malloc(&p, std::max<uint64_t>(cfg.budget_bytes, 256)));
        d_arena = (uint8_t *)p;
        // group staging: a batch of the largest group-bys, at most 8 Mi groups (128 MiB) per slot
        const uint64_t gcap = std::min<uint64_t>((uint64_t)cfg.max_batch * max_groups, 8ull << 20);
        slots.resize(cfg.max_inflight);
        for (Slot &s : slots) {
            gpu(mgv_stream_create(&s.stream));
            gpu(mgv_malloc(&p, desc_bytes));
            s.d_desc = (uint8_t *)p;
            gpu(mgv_malloc(&p, 2 * (size_t)cfg.max_batch * sizeof(mgv_result)));
            s.d_res = (mgv_result *)p;
            gpu(mgv_malloc(&p, 256));
            s.d_counter = (uint32_t *)p;
            gpu(mgv_host_malloc(&p, desc_bytes));
            s.h_desc = (uint8_t *)p;
            gpu(mgv_host_malloc(&p, 2 * (size_t)cfg.max_batch * sizeof(mgv_result)));
            s.h_res = (mgv_result *)p;
            s.group_cap = std::max<uint64_t>(gcap, max_groups);
            if (s.group_cap) {
                gpu(mgv_host_malloc(&p, s.group_cap * 16));
                s.h_groups = (uint64_t *)p;
            }
        }
    }

    // Builds the batch's descriptors in the slot's pinned staging and issues it.
    void issue(Slot &s) {
        mgv_query *bq = (mgv_query *)s.h_desc;
        mgv_query *mq = bq + cfg.max_batch;
        uint64_t *ids = (uint64_t *)(s.h_desc + ids_at);
        mgv_work_item *items = (mgv_work_item *)(s.h_desc + items_at);
        for (size_t i = 0; i < s.q.size(); ++i) ids[i] = s.q[i].req;
        s.nbuild = s.nmain = 0;
        uint32_t join = 0;   // build result slot: max_batch + the join's index in the batch
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            mgv_query m;
            memset(&m, 0, sizeof(m));
            if (r.kind == MGV_SERVE_JOIN_SUM) {
                mgv_query b;
                memset(&b, 0, sizeof(b));
                b.kind = MGV_Q_BUILD;
                b.key = r.bkey;
                b.val = r.bval;
                b.n = r.bn;
                b.slots = (uint64_t *)(d_arena + a.off);
                b.nbuckets = a.nbuckets;
                b.res = s.d_res + cfg.max_batch + join++;
                b.payload_mode = MGV_PAYLOAD_COLUMN;
                if (!a.ws) bq[s.nbuild++] = b;   // large builds run through LDS slices below
                m.kind = MGV_Q_PROBE_SUM;
                m.bound = r.bound;
                m.pred = r.pred;
                m.slots = b.slots;
                m.nbuckets = a.nbuckets;
            } else {
                if (a.ws) continue;   // many groups: partitioned group-by, issued below
                m.kind = MGV_Q_GROUPBY;
                m.ngroups = r.ngroups;
                m.gcnt = (uint64_t *)(d_arena + a.off);
                m.gsum = (int64_t *)(d_arena + a.off + (uint64_t)r.ngroups * 8);
            }
            m.key = r.key;
            m.val = r.val;
            m.n = r.n;
            m.res = s.d_res + i;   // result slot i belongs to the batch's i-th request
            mq[s.nmain++] = m;
        }
        uint64_t nb = 0, nm = 0;
        if (s.nbuild) plan(bq, s.nbuild, items, items_cap, &nb);
        if (s.nmain) plan(mq, s.nmain, items + nb, items_cap - nb, &nm);
        if (nb + nm > items_cap) throw mgv::error(MGV_E_CAPACITY, "serve: work-item staging too small");
        s.nbitems = nb;
        s.nmitems = nm;
        // one copy of every descriptor, then clears, builds, probes / group-bys, result copies
        const size_t bytes = items_at + (size_t)(nb + nm) * sizeof(mgv_work_item);
        gpu(mgv_memcpy_async(s.d_desc, s.h_desc, bytes, s.stream));
        gpu(mgv_result_clear(s.d_res, 2 * cfg.max_batch, s.stream));
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            if (r.kind == MGV_SERVE_JOIN_SUM) gpu(mgv_ht_clear((uint64_t *)(d_arena + a.off), a.nbuckets, s.stream));
            else gpu(mgv_fill_u64((uint64_t *)(d_arena + a.off), 2ull * r.ngroups, 0, s.stream));
        }
        const mgv_query *dbq = (const mgv_query *)s.d_desc;
        const mgv_query *dmq = dbq + cfg.max_batch;
        const mgv_work_item *di = (const mgv_work_item *)(s.d_desc + items_at);
        join = 0;
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            if (r.kind != MGV_SERVE_JOIN_SUM) continue;
            const uint32_t j = join++;
            if (!a.ws) continue;
            gpu(mgv_hashjoin_build_lds((uint64_t *)(d_arena + a.off), a.nbuckets, r.bkey, r.bval, MGV_PAYLOAD_COLUMN,
                                       nullptr, 0, r.bn, 0, kLdsBuildFanout, d_arena + a.off + (a.bytes - a.ws), a.ws,
                                       s.d_res + cfg.max_batch + j, mgv_launch{0, 1024, -1, MGV_LAUNCH_SCATTER(2u)},
                                       s.stream));
        }
        if (nb) gpu(mgv_batch_run(dbq, di, nb, s.d_counter, cfg.launch, s.stream));
        if (nm) gpu(mgv_batch_run(dmq, di + nb, nm, s.d_counter + 32, cfg.launch, s.stream));
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            if (r.kind != MGV_SERVE_GROUPBY || !a.ws) continue;
            uint8_t *g = d_arena + a.off;
            const uint64_t garr = align_up((uint64_t)r.ngroups * 16, 256);
            gpu(mgv_groupby_partitioned(r.key, r.val, r.n, (uint64_t *)g, (int64_t *)(g + (uint64_t)r.ngroups * 8),
                                        r.ngroups, g + garr, a.ws, s.d_res + i, mgv_launch{0, 1024, -1, 0},
                                        s.stream));
        }
        gpu(mgv_memcpy_async(s.h_res, s.d_res + cfg.max_batch, (size_t)cfg.max_batch * sizeof(mgv_result),
                             s.stream));
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            if (r.kind == MGV_SERVE_GROUPBY)
                gpu(mgv_memcpy_async(s.h_groups + 2 * a.group_stage, d_arena + a.off, (size_t)r.ngroups * 16,
                                     s.stream));
        }
        // gpu_to_cpu: every request's {id, result slot, result} on the ring,
        // after the copies above (stream order)
        gpu(mgv_ring_emit_results(ring, (const uint64_t *)(s.d_desc + ids_at), s.d_res, (uint32_t)s.q.size(),
                                  s.stream));
        s.pending = (uint32_t)s.q.size();
        s.busy = true;
    }

    // A request's completion came off the ring: its result, and the batch's
    // retirement once it was the batch's last.
    void complete(const mgv_ring_entry &e, double now) {
        if (e.id >= n) throw mgv::error(MGV_E_INVALID, "serve: completion of an unknown request");
        Slot &s = slots[slot_of[e.id]];
        if (!s.busy || e.value >= s.q.size() || s.q[e.value].req != e.id)
            throw mgv::error(MGV_E_INVALID, "serve: completion does not match its batch");
        s.ring_res.resize(s.q.size());
        s.ring_res[e.value] = e.res;
        out[e.id].done_s = now;
        ++ring_results;
        if (--s.pending == 0) retire(s);
    }

    void retire(Slot &s) {
        uint32_t b = 0;
        for (size_t i = 0; i < s.q.size(); ++i) {
            const Admitted &a = s.q[i];
            const mgv_serve_request &r = R[a.req];
            mgv_serve_result &o = out[a.req];
            const mgv_result &m = s.ring_res[i];
            const double now = o.done_s;
            o.latency_s = now - r.arrival_s;
            o.batch = s.batch;
            o.footprint = align_up(a.bytes, 256);
            if (r.kind == MGV_SERVE_JOIN_SUM) {
                o.count = m.count;
                o.sum = m.sum;
                o.digest = 0;
                o.errors = (uint32_t)s.h_res[b].errors;
                ++b;
            } else {
                const uint64_t *cnt = s.h_groups + 2 * a.group_stage;
                const int64_t *sum = (const int64_t *)(cnt + r.ngroups);
                uint64_t h = 0x9E3779B97F4A7C15ull, c = 0;
                int64_t t = 0;
                for (uint32_t g = 0; g < r.ngroups; ++g) {
                    h = mix64(h ^ cnt[g]);
                    h = mix64(h ^ (uint64_t)sum[g]);
                    c += cnt[g];
                    t += sum[g];
                }
                o.count = c;
                o.sum = t;
                o.digest = h;
                o.errors = (uint32_t)m.errors;
            }
            arena.release(a.off, a.bytes);
        }
        s.q.clear();
        s.busy = false;
    }

    void run(mgv_serve_stats *st) {
        setup();
        std::deque<uint32_t> waiting;
        uint32_t next = 0, done = 0, batches = 0, max_inflight_seen = 0;
        uint64_t batched_queries = 0;
        const auto t0 = clock_t_::now();
        auto now_s = [&] { return std::chrono::duration<double>(clock_t_::now() - t0).count(); };
        while (done < n) {
            double now = now_s();
            while (next < n && R[next].arrival_s <= now) waiting.push_back(next++);
            // completed requests off the ring; a batch retires with its last
            for (;;) {
                mgv_ring_entry e;
                const int q = mgv_ring_poll(ring, &e);
                if (q == 1) break;
                gpu(q);
                complete(e, now_s());
                ++done;
            }
            // admit into free slots
            bool issued = false;
            for (Slot &s : slots) {
                if (s.busy || waiting.empty()) continue;
                uint64_t items = 0, gstage = 0, rows = 0;
                while (!waiting.empty() && s.q.size() < cfg.max_batch) {
                    const uint32_t i = waiting.front();
                    const mgv_serve_request &r = R[i];
                    uint32_t nb = 0;
                    uint64_t ws = 0;
                    const uint64_t bytes = footprint(r, &nb, &ws);
                    const uint64_t it = items_of(r.n) + (r.kind == MGV_SERVE_JOIN_SUM ? items_of(r.bn) : 0);
                    const uint64_t gs = r.kind == MGV_SERVE_GROUPBY ? r.ngroups : 0;
                    const uint64_t rr = r.n + (r.kind == MGV_SERVE_JOIN_SUM ? r.bn : 0);
                    if (items + it > items_cap || gstage + gs > s.group_cap) break;
                    if (!s.q.empty() && rows + rr > kBatchRows) break;   // the next batch takes it
                    uint64_t off = 0;
                    if (!arena.alloc(bytes, &off)) break;   // FIFO: wait for memory
                    s.q.push_back(Admitted{i, off, bytes, nb, gstage, ws});
                    slot_of[i] = (uint32_t)(&s - slots.data());
                    out[i].admit_s = now_s();
                    items += it;
                    gstage += gs;
                    rows += rr;
                    waiting.pop_front();
                }
                if (s.q.empty()) break;   // the head of the queue does not fit yet
                s.batch = batches++;
                batched_queries += s.q.size();
                issue(s);
                issued = true;
            }
            uint32_t inflight = 0;
            for (Slot &s : slots) inflight += s.busy;
            max_inflight_seen = std::max(max_inflight_seen, inflight);
            if (!issued) {
                // idle: wait for the next arrival or completion
                const double wait = next < n ? R[next].arrival_s - now_s() : 1e-4;
                if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds(
                                  (int64_t)std::min(50.0, std::max(5.0, wait * 1e6))));
            }
        }
        const double end = now_s();
        if (st) {
            std::vector<double> lat(n);
            double mean = 0;
            for (uint32_t i = 0; i < n; ++i) {
                lat[i] = out[i].latency_s;
                mean += lat[i];
            }
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) {
                if (lat.empty()) return 0.0;
                const size_t k = (size_t)std::ceil(p * lat.size()) - 1;
                return lat[std::min(k, lat.size() - 1)];
            };
            st->elapsed_s = n ? end - R[0].arrival_s : 0;
            st->queries_per_s = st->elapsed_s > 0 ? n / st->elapsed_s : 0;
            st->p50_s = pct(0.50);
            st->p99_s = pct(0.99);
            st->max_s = lat.empty() ? 0 : lat.back();
            st->mean_s = n ? mean / n : 0;
            st->batches = batches;
            st->max_inflight_seen = max_inflight_seen;
            st->peak_bytes = arena.peak();
            st->mean_batch = batches ? (double)batched_queries / batches : 0;
            st->ring_results = ring_results;
        }
    }
};

}  // namespace

namespace mgvhost {

// Runs the loop; throws mgv::error (the C entry point is in drivers.cpp).
void serve_run(const mgv_serve_request *requests, uint32_t n, const mgv_serve_config *config,
               mgv_serve_result *results, mgv_serve_stats *stats) {
    if ((!requests && n) || !config || (!results && n)) throw mgv::error(MGV_E_INVALID, "null argument");
    if (n) memset(results, 0, sizeof(mgv_serve_result) * n);
    if (stats) memset(stats, 0, sizeof(*stats));
    Server s(requests, n, *config, results);
    s.run(stats);
}

}  // namespace mgvhost
