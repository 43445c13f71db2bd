/*
 * mgv.h -- C ABI of libmgv.so, the MI355X (gfx950) implementation of the
 * vector-at-a-time select -> hash-join -> aggregation operator path of
 * chapeiro/multigpu-vector.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns 0 on
 * success, otherwise a HIP error code or one of the MGV_E* codes below, and
 * records a message for mgv_last_error_string().  Nothing here calls exit().
 * Device pointers come from mgv_malloc; every launch takes an explicit stream
 * (NULL = the library's default stream of the current device).  All of the
 * library's streams, the default one included, are non-blocking: work on two
 * different streams is ordered only through events or synchronisation.
 * Calls are thread-safe for distinct streams.
 *
 * Each entry point names the reference interface it replaces (path:line,
 * relative to the reference tree).  The reference has no C ABI of its own:
 * its boundary is the C++ operator concept dispatched through
 * d_operator_t / h_operator_t (src/operators/operator.cuh:42-124) plus the
 * buffer_manager / buffer_pool contract; the C++ drop-in layer in
 * multigpu-vector_amd/host/ re-exposes that concept on top of this ABI.
 */
#ifndef MGV_H_
#define MGV_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGV_ABI_VERSION 1

/* Error codes beyond HIP's own (hipError_t values are < 10000). */
#define MGV_OK 0
#define MGV_E_INVALID 10001     /* bad argument / shape */
#define MGV_E_NOT_INIT 10002    /* mgv_init not called on this thread's device */
#define MGV_E_CAPACITY 10003    /* output / pool capacity exceeded */
#define MGV_E_ALIGN 10004       /* pointer alignment requirement violated */
#define MGV_E_RESERVED 10005    /* build row equal to the reserved empty slot */
#define MGV_E_NOGPU 10006       /* no HIP device */

typedef void *mgv_stream_t;   /* hipStream_t */
typedef void *mgv_event_t;    /* hipEvent_t */
typedef void *mgv_graph_t;    /* hipGraphExec_t (+ graph) */

/* ------------------------------------------------------------ runtime */

/* Select `device`, create the library's default stream and read device
 * properties.  Replaces the device part of buffer_manager<int32_t>::init
 * (src/buffer_manager.cu:38-211) and the gpu()/exit error convention
 * (src/common.cuh:12-23). */
int mgv_init(int device);
int mgv_shutdown(void);
int mgv_abi_version(void);
const char *mgv_last_error_string(void);

typedef struct mgv_device_props {
    char name[64];
    char arch[32];            /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t cus;              /* compute units */
    int32_t wave_size;        /* 64 */
    int64_t lds_per_block;    /* bytes */
    int64_t hbm_bytes;
    int64_t l2_bytes;
    int32_t clock_khz;
    int32_t mem_clock_khz;
    int32_t mem_bus_width;
    int32_t max_threads_per_block;
} mgv_device_props;
int mgv_get_device_props(mgv_device_props *out);

/* Memory: device, pinned host.  cudaMalloc / cudaMallocHost call sites of
 * src/buffer.cuh:34-41, src/buffer_manager.cu:98-176. */
int mgv_malloc(void **ptr, size_t bytes);
int mgv_free(void *ptr);
int mgv_host_malloc(void **ptr, size_t bytes);
int mgv_host_free(void *ptr);
int mgv_mem_get_info(size_t *free_bytes, size_t *total_bytes);

/* Copies (direction inferred from the pointers).  Replaces
 * buffer_manager::overwrite (src/buffer_manager.cu:214-218) and the
 * buffer_inspector copies (src/buffer.cuh:183-217). */
int mgv_memcpy(void *dst, const void *src, size_t bytes);
int mgv_memcpy_async(void *dst, const void *src, size_t bytes, mgv_stream_t s);

/* Fill with a 32-/64-bit pattern by kernel (graph-safe; no memset nodes). */
int mgv_fill_u32(uint32_t *dst, uint64_t n, uint32_t value, mgv_stream_t s);
int mgv_fill_u64(uint64_t *dst, uint64_t n, uint64_t value, mgv_stream_t s);

/* Streams / events / sync (the reference's per-consumer streams,
 * src/operators/exchange.cu:293-318).  Streams are created at default
 * priority. */
int mgv_stream_create(mgv_stream_t *s);
int mgv_stream_destroy(mgv_stream_t s);
int mgv_stream_synchronize(mgv_stream_t s);
int mgv_device_synchronize(void);
int mgv_event_create(mgv_event_t *e);
int mgv_event_destroy(mgv_event_t e);
int mgv_event_record(mgv_event_t e, mgv_stream_t s);
int mgv_event_synchronize(mgv_event_t e);
int mgv_event_query(mgv_event_t e);   /* 0 = complete, 1 = pending, else error */
int mgv_event_elapsed_ms(float *ms, mgv_event_t start, mgv_event_t stop);
int mgv_stream_wait_event(mgv_stream_t s, mgv_event_t e);

/* Graphs: capture a stream into an executable graph and replay it. */
int mgv_graph_begin_capture(mgv_stream_t s);
int mgv_graph_end_capture(mgv_stream_t s, mgv_graph_t *g);
int mgv_graph_launch(mgv_graph_t g, mgv_stream_t s);
int mgv_graph_destroy(mgv_graph_t g);

/* ------------------------------------------------------------ launch shape */

/* Launch shape of a pipeline.  Mirrors launch_conf{gridDim, blockDim,
 * shared_mem, device} (src/common.cuh:183-209): grid 0 = persistent grid
 * sized from the CU count; block in {64,128,256,512,1024}; compact selects
 * LDS stream compaction (1) or predication (0) between select and its parent
 * (-1 = the operator's default). */
typedef struct mgv_launch {
    uint32_t grid;
    uint32_t block;
    int32_t compact;
    uint32_t reserved;        /* flags: MGV_LAUNCH_* */
} mgv_launch;

/* Launch flag: the probe-sum pipelines use the wave-cooperative bucket scan
 * (4 lanes per row) instead of one row per lane. */
#define MGV_LAUNCH_COOP_PROBE 1u
/* Workgroup shape of the partitioning scatter (bits 8-11 of the flags) used
 * by the partitioned probes and group-by: threads x rows per LDS tile. */
#define MGV_SCATTER_256x4096 0u
#define MGV_SCATTER_512x8192 1u
#define MGV_SCATTER_1024x8192 2u
#define MGV_SCATTER_1024x4096 3u
#define MGV_LAUNCH_SCATTER(shape) ((uint32_t)(shape) << 8)
/* Launch flag of mgv_hashjoin_probe_lds: partition with count passes and
 * exact partition sizes instead of the default count-free partitioning into
 * fixed-capacity regions (same results; kept for comparison). */
#define MGV_LAUNCH_COUNTED_PART (1u << 12)
/* Launch flag of mgv_hashjoin_probe_lds / mgv_hashjoin_probe_groupby_lds:
 * hold the slices as 32-bit key hashes instead of the 16-bit slot tags used
 * by default for power-of-two tables of at least 2^18 buckets (same
 * results; kept for comparison). */
#define MGV_LAUNCH_KEY_SLICES (1u << 13)
/* Launch flag of mgv_batch_run_lds: the tables of the queries it builds
 * through LDS slices (regions word F != 0) are written whole, every slice
 * from an empty start, whatever they held: they need no clear before the
 * call.  The tables of its other build queries are inserted into as they
 * stand (clear them first, as for mgv_batch_run). */
#define MGV_LAUNCH_FRESH_TABLES (1u << 14)

/* ------------------------------------------------------------ datagen */

/* Counter-based Philox4x32-10 columns; identical to the oracle's generator. */
int mgv_gen_bits(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t bits, mgv_stream_t s);
int mgv_gen_range(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t range, mgv_stream_t s);
int mgv_gen_perm(int32_t *dst, uint64_t n, uint32_t range, uint64_t seed, uint32_t pstream, mgv_stream_t s);
int mgv_gen_iota(int32_t *dst, uint64_t n, int32_t base, mgv_stream_t s);
/* Zipf(s) thresholds (host-computed, uploaded to *table, nranks+... entries). */
int mgv_zipf_table_create(uint64_t **table, uint32_t nranks, double s_exp);
int mgv_gen_zipf(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, const uint64_t *table,
                 uint32_t nranks, uint64_t pseed, uint32_t pstream, mgv_stream_t s);

/* ------------------------------------------------------------ results */

/* Scalar result of an aggregating pipeline, device resident, accumulated with
 * atomics (clear before use).  count/sum replace the int32 `res` of
 * aggregation (src/operators/aggregation.cuh:12) and hashjoin
 * (src/operators/hashjoin.cuh:41) with exact 64-bit values. */
typedef struct mgv_result {
    uint64_t count;
    int64_t sum;
    uint64_t digest_sum;      /* sum of mix64(r_row << 32 | s_payload) over join matches */
    uint64_t digest_xor;      /* xor of the same */
    uint64_t errors;          /* rejected rows (reserved build pattern) */
    uint64_t overflow;        /* rows that did not fit an output buffer */
    uint64_t reserved[2];
} mgv_result;
int mgv_result_clear(mgv_result *r, uint32_t n, mgv_stream_t s);

/* ------------------------------------------------------------ hash table */

/* Open-addressing hash table: nbuckets buckets of 8 slots of 8 bytes
 * {key:int32, payload:int32} (64-byte buckets), bucket = (fmix32(key) *
 * nbuckets) >> 32, bucketised linear probing.  The slot pattern
 * {key=-1, payload=-1} marks an empty slot and is rejected by the build.
 * Replaces the chained first[]/next[] table of hashjoin::hashjoin
 * (src/operators/hashjoin.cu:63-75). */
#define MGV_SLOTS_PER_BUCKET 8
int mgv_ht_plan(uint64_t nkeys, double load_factor, uint32_t *nbuckets, uint64_t *bytes);
int mgv_ht_clear(uint64_t *slots, uint32_t nbuckets, mgv_stream_t s);

/* payload modes of the build */
#define MGV_PAYLOAD_NONE 0     /* payload 0 (the reference's keys-only table) */
#define MGV_PAYLOAD_COLUMN 1   /* payload[i] */
#define MGV_PAYLOAD_ROWID 2    /* row_base + i (for join-pair digests) */

/* Build: insert every row i (optionally only rows with pred[i] <= bound).
 * Replaces hashjoin_builder::consume_warp (src/operators/hashjoin.cu:41-60).
 * Rejected rows are counted in res->errors when res != NULL. */
int mgv_hashjoin_build(uint64_t *slots, uint32_t nbuckets, const int32_t *keys, const int32_t *payload,
                       int payload_mode, const int32_t *pred, int32_t bound, uint64_t n, int64_t row_base,
                       mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* Probe flags */
#define MGV_PROBE_DIGEST 1     /* also accumulate the match-pair digest (payload = s row id) */

/* Probe + aggregate: for every probe row r (optionally pred[r] <= bound),
 * for every slot with key == keys[r]: count += 1, sum += vals[r] + payload
 * (vals may be NULL).  Replaces hashjoin::consume_warp + at_close
 * (src/operators/hashjoin.cu:88-96,135-182) fused with aggregation
 * (src/operators/aggregation.cu:27-66). */
int mgv_hashjoin_probe(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                       const int32_t *keys, const int32_t *vals, uint64_t n, int64_t row_base, uint32_t flags,
                       mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* Probe + group-by: group g = payload of the matching slot (g < ngroups),
 * cnt[g] += 1, sum[g] += vals[r].  No reference counterpart for group-by;
 * the probe part replaces src/operators/hashjoin.cu:135-182. */
int mgv_hashjoin_probe_groupby(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                               const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt,
                               int64_t *gsum, uint32_t ngroups, mgv_launch launch, mgv_stream_t s);

/* Probe + materialise the match multiset as (r_row, payload) pairs appended
 * to out_r/out_p (capacity cap); res->count = total matches, res->overflow =
 * pairs beyond cap.  Replaces the hashjoin -> gpu_to_cpu chain of
 * src/main9.cu:326-367. */
int mgv_hashjoin_probe_materialize(const uint64_t *slots, uint32_t nbuckets, const int32_t *keys, uint64_t n,
                                   int64_t row_base, int32_t *out_r, int32_t *out_p, uint64_t cap,
                                   mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* ------------------------------------------------------------ buffer pool */

/* Device buffer pool: nbuffers buffers of `rows` int32 (rows a multiple of
 * 256) in one device arena, each with a device-visible count, and a
 * lock-free stack of free buffers (64-bit {index, ABA tag} head, one-lane
 * operations) with the reference's producer / available counters.  Replaces
 * buffer_pool<T> (src/buffer_pool.cuh:35-170), the device half of
 * buffer_manager<int32_t> (src/buffer_manager.cuh:40-181) and the device
 * stacks (src/lockfree_stack.cuh:19-208, src/threadsafe_device_stack.cuh:
 * 19-87). */
typedef struct mgv_pool mgv_pool;
#define MGV_POOL_END (-1)     /* no producer registered and no buffer available (end of stream) */
#define MGV_POOL_RETRY (-2)   /* 100 attempts failed; the reference's (buffer_t *)1 sentinel */
int mgv_pool_create(uint32_t nbuffers, uint32_t rows, mgv_pool **pool);
/* The same pool with its buffers in pinned coherent host memory: kernels
 * write rows over PCIe and the host reads them in place (mgv_pool_buffer
 * returns host pointers), so handing a full buffer to the host needs no
 * copy engine or kernel.  For output bound to the host (gpu_to_cpu). */
int mgv_pool_create_host(uint32_t nbuffers, uint32_t rows, mgv_pool **pool);
int mgv_pool_destroy(mgv_pool *pool);
/* acquire_buffer_blocked_unsafe (src/buffer_pool.cuh:83-96) run by one lane
 * on the device; *buffer = index, MGV_POOL_END or MGV_POOL_RETRY. */
int mgv_pool_acquire(mgv_pool *pool, int32_t *buffer, mgv_stream_t s);
/* release_buffer (src/buffer_pool.cuh:125-130); the buffer count is reset
 * when it is next acquired. */
int mgv_pool_release(mgv_pool *pool, int32_t buffer, mgv_stream_t s);
/* register_producer / unregister_producer (src/buffer_pool.cuh:136-165). */
int mgv_pool_register_producer(mgv_pool *pool, mgv_stream_t s);
int mgv_pool_unregister_producer(mgv_pool *pool, mgv_stream_t s);
/* Device pointers of buffer `buffer`: its rows and its row count. */
int mgv_pool_buffer(mgv_pool *pool, int32_t buffer, int32_t **data, uint32_t **count);
int mgv_pool_info(mgv_pool *pool, uint32_t *nbuffers, uint32_t *rows);
/* available / producers counters and the number of buffers on the stack
 * (the last only meaningful when no kernel uses the pool). */
int mgv_pool_stats(mgv_pool *pool, uint32_t *available, uint32_t *producers, uint32_t *stacked, mgv_stream_t s);
/* Concurrency check of the stack: `blocks` x 4 waves each acquire / verify
 * exclusive ownership / release `iters` times; ops4 (device, 4 x uint64,
 * accumulated): pairs done, ownership violations, RETRY results, END results. */
int mgv_pool_stress(mgv_pool *pool, uint32_t blocks, uint32_t iters, uint64_t *ops4, mgv_stream_t s);
/* Append {val[i] : pred[i] <= bound} to pool buffer `buffer` with one
 * atomic reservation per wave (buffer::try_write / try_partial_final_write,
 * src/buffer.cuh:92-134); rows past the buffer capacity are counted but not
 * written. */
int mgv_select_to_buffer(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_pool *pool,
                         int32_t buffer, mgv_stream_t s);

/* ------------------------------------------------------------ result channel */

/* Device -> host result channel: pinned, coherent slots {flag, mgv_result}.
 * emit copies a device result into a slot and publishes it with a
 * system-scope release; poll returns 0 and the result once the slot is set
 * (and clears it), 1 while it is not.  The reduced gpu_to_cpu ring
 * (src/operators/gpu_to_cpu.cu:42-58,108-124,149-165). */
typedef struct mgv_channel mgv_channel;
int mgv_channel_create(uint32_t nslots, mgv_channel **channel);
int mgv_channel_destroy(mgv_channel *channel);
int mgv_channel_emit(mgv_channel *channel, uint32_t slot, const mgv_result *res, uint32_t seq, mgv_stream_t s);
int mgv_channel_poll(mgv_channel *channel, uint32_t slot, mgv_result *out, uint32_t *seq);

/* ------------------------------------------------------------ completion ring */

/* Device -> host ring of completions: slots in pinned coherent host memory,
 * each claimed by a ticket (one atomic on a device counter) and published
 * with a system-scope release; the host takes them in ticket order.  Used
 * for full output buffers (id = pool buffer, value = rows) and completed
 * queries (id = query id, value = result slot, res = its result).  A slot
 * the host has not taken yet makes the push fail instead of wait: the entry
 * is counted as lost and mgv_ring_poll reports MGV_E_CAPACITY, so size the
 * ring for the most entries that can be outstanding.  Replaces gpu_to_cpu's
 * ring (src/operators/gpu_to_cpu.cu:42-58 throw2cpu, 149-165 catcher). */
typedef struct mgv_ring_entry {
    uint64_t id;
    uint64_t value;
    mgv_result res;
} mgv_ring_entry;
#define MGV_RING_EOF (~0ull)   /* id of the entry mgv_stream_out_close publishes last */
typedef struct mgv_ring mgv_ring;
int mgv_ring_create(uint32_t nslots, mgv_ring **ring);
int mgv_ring_destroy(mgv_ring *ring);
/* One entry {id, value, *res (zeros when res is NULL)} pushed by a kernel on
 * stream s, after the work before it on s. */
int mgv_ring_emit(mgv_ring *ring, uint64_t id, uint64_t value, const mgv_result *res, mgv_stream_t s);
/* n entries {ids[i], i, res[i]} (ids and res device arrays). */
int mgv_ring_emit_results(mgv_ring *ring, const uint64_t *ids, const mgv_result *res, uint32_t n, mgv_stream_t s);
/* Host: 0 and the oldest entry, 1 when none is published yet; MGV_E_CAPACITY
 * once an entry was lost.  One host thread polls a ring. */
int mgv_ring_poll(mgv_ring *ring, mgv_ring_entry *out);

/* ------------------------------------------------------------ rotating output buffers */

/* Hands a pool buffer back from the host without a kernel: the index goes
 * into the pool's return ring in pinned memory, from which the device takes
 * it when its stack is empty.  Safe while kernels run. */
int mgv_pool_release_host(mgv_pool *pool, int32_t buffer);

/* Output of a streaming select: rows are appended to the current pool
 * buffer with one reservation per wave (buffer::try_write /
 * try_partial_final_write, src/buffer.cuh:92-134); a wave whose reservation
 * does not fit replaces the buffer with a fresh one from the pool
 * (buffer_manager::acquire_buffer_blocked_to, src/buffer_manager.cuh:
 * 147-176: one lane, a CAS of a tagged {tag, index} word).  A replaced
 * buffer is published on the ring as {buffer, rows} once every row reserved
 * in it is written; the host copies it out and hands it back with
 * mgv_pool_release_host.  When no buffer comes back within 10 s the wave's
 * rows are dropped and counted (stats), never waited for longer.  The ring
 * needs nbuffers + 2 slots.  The thread that hands buffers back must not
 * wait for the device while the kernel runs (hipFree, hipStreamCreate and
 * hipDeviceSynchronize wait for every queue): the kernel waits for it in
 * turn, and its rows are dropped when the 10 s pass. */
typedef struct mgv_stream_out mgv_stream_out;
int mgv_stream_out_create(mgv_pool *pool, mgv_ring *ring, mgv_stream_out **out);
int mgv_stream_out_destroy(mgv_stream_out *out);
/* Starts a new stream of rows (no current buffer; counters zero). */
int mgv_stream_out_open(mgv_stream_out *out, mgv_stream_t s);
/* After the last append on s: retires the current buffer and publishes
 * {MGV_RING_EOF, rows handed over, res.overflow = rows dropped}
 * (gpu_to_cpu::at_close, src/operators/gpu_to_cpu.cu:108-124). */
int mgv_stream_out_close(mgv_stream_out *out, mgv_stream_t s);
/* Appends {val[i] : pred[i] <= bound} (unstable_select -> gpu_to_cpu,
 * src/operators/select3.cu:19-85 with gpu_to_cpu::consume_warp,
 * src/operators/gpu_to_cpu.cu:61-105).  val may equal pred. */
int mgv_select_stream(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_stream_out *out,
                      mgv_launch launch, mgv_stream_t s);

/* ------------------------------------------------------------ fused query batches */

/* Query kinds of a fused batch. */
#define MGV_Q_SELECT_SUM 1      /* COUNT, SUM(val) WHERE pred <= bound (pred NULL = all rows) */
#define MGV_Q_PROBE_SUM 2       /* [select] -> probe -> COUNT, SUM(val + payload) [+ digest] */
#define MGV_Q_PROBE_GROUPBY 3   /* [select] -> probe -> GROUP BY payload: COUNT, SUM(val) */
#define MGV_Q_GROUPBY 4         /* GROUP BY key: COUNT, SUM(val) */
#define MGV_Q_BUILD 5           /* [select] -> build (payload_mode) */

/* One query of a batch, device resident.  Row ranges of a query are its
 * vectors; all pointers are device pointers with 16-byte aligned columns. */
typedef struct mgv_query {
    uint32_t kind;
    int32_t bound;            /* select: pred <= bound; ignored when pred is NULL */
    const int32_t *pred;
    const int32_t *key;
    const int32_t *val;       /* summed / carried column; BUILD payload column */
    uint64_t n;               /* rows */
    uint64_t *slots;          /* hash table (PROBE_*, BUILD) */
    uint32_t nbuckets;
    uint32_t ngroups;
    uint64_t *gcnt;           /* group-by outputs */
    int64_t *gsum;
    mgv_result *res;          /* COUNT/SUM; errors of builds and group-bys */
    int32_t payload_mode;     /* BUILD: MGV_PAYLOAD_* */
    uint32_t flags;           /* PROBE_SUM: MGV_PROBE_DIGEST */
    int64_t row_base;         /* row id of row 0 (digest / ROWID payload) */
} mgv_query;

/* One vector of one query: rows [lo, lo + rows). */
typedef struct mgv_work_item {
    uint32_t query;
    uint32_t rows;
    uint64_t lo;
} mgv_work_item;

/* Host-side: cut every query (host copy of the table, same device pointers)
 * into vectors of vector_rows rows (a multiple of 256), in query order;
 * *nitems = number of items (also when items is NULL).  Validates the
 * queries.  Replaces the generator's slicing of a column into buffers
 * (src/operators/generators.cu:55-90). */
int mgv_batch_plan(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, mgv_work_item *items,
                   uint64_t cap, uint64_t *nitems);

/* One persistent launch drains the work items of a batch (device arrays):
 * each wave claims a vector with one atomic on *counter (reset by this call
 * with a fill kernel, so the call can be captured into a graph) and runs its
 * query's chain.  Builds must be in an earlier call than the probes that
 * use their tables.  launch.block: 64, 128, 256, 512 or 1024 threads
 * (0 = 256).  Replaces one launch_consume_pipeline per vector
 * (src/operators/exchange.cu:124-194,196-291) with one launch per batch. */
int mgv_batch_run(const mgv_query *queries, const mgv_work_item *items, uint64_t nitems, uint32_t *counter,
                  mgv_launch launch, mgv_stream_t s);

/* XCD-sliced batch probe (same per-query results as mgv_batch_run).  A
 * probe's bucket reads that miss the L2 fetch a line each from HBM, and every
 * XCD's L2 holds the whole table when every XCD probes it; here the selected
 * rows of each large PROBE_SUM query are first scattered by the top bits
 * of fmix32(key) into nregions = 8, 16 or 32 regions (0: 32; the rows whose
 * home buckets lie in one nregions-th of the table), then one persistent
 * launch lets the waves of XCD x probe regions x, x + 8, ... of every query,
 * one after another, so each L2 holds the part being probed and no line is
 * cached twice.  Rows past a full region (skewed keys) are probed directly.
 * mgv_batch_plan_sliced (host): validates the queries as mgv_batch_plan;
 * queries of kind PROBE_SUM without MGV_PROBE_DIGEST and with n >= min_rows
 * are sliced; fills items (capacity cap; *plan tells the count, also when
 * items is NULL), regions (4 words per query: first region row, rows per
 * region, regions (0 for a query that is not sliced), 0) and *plan.  Items: scatter tiles
 * [0, nscatter), then probe sections x = 0..7 (regions x, x + 8, ... of every
 * sliced query, vectors of vector_rows rows over each region's capacity;
 * the item's query field holds the query index | region << 24) and section
 * 8 (the other queries' vectors).  nqueries < 2^24.
 * mgv_batch_run_sliced: device copies of queries, items and regions; the
 * workspace (plan->workspace_bytes, 16-byte aligned) holds the claim
 * counters, fill counts and region columns (8 bytes per row of nregions
 * regions of about n / nregions rows each per sliced query); clears them
 * with a fill kernel,
 * so the call can be captured into a graph.  launch.block / compact as for
 * mgv_batch_run.  No reference counterpart (the reference's per-vector
 * probe, src/operators/hashjoin.cu:135-182, reads the whole table from every
 * SM). */
typedef struct mgv_sliced_plan {
    uint64_t nitems;             /* all items */
    uint64_t nscatter;           /* scatter tiles: items [0, nscatter) */
    uint64_t section[10];        /* probe section s: items [section[s], section[s + 1]) */
    uint64_t region_rows;        /* rows of all regions */
    uint64_t workspace_bytes;
    uint32_t nqueries;
    uint32_t nregions;           /* regions per sliced query: 8, 16 or 32 */
} mgv_sliced_plan;
int mgv_batch_plan_sliced(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, uint64_t min_rows,
                          uint32_t nregions, mgv_work_item *items, uint64_t cap, uint64_t *regions,
                          mgv_sliced_plan *plan);
int mgv_batch_run_sliced(const mgv_query *queries, const mgv_work_item *items, const uint64_t *regions,
                         const mgv_sliced_plan *plan, void *ws, uint64_t ws_bytes, mgv_launch launch, mgv_stream_t s);

/* LDS batch probe and build (same per-query results and table contents as
 * mgv_batch_run): the PROBE_SUM queries (no MGV_PROBE_DIGEST, n >= min_rows)
 * whose table splits into at most 256 power-of-two slices of at most 2048
 * buckets, and the BUILD queries (not MGV_PAYLOAD_ROWID, n >= min_rows) whose
 * table splits into at most 256 slices of at most 1024 buckets, have their
 * selected rows scattered (hashed keys) into one region per slice; one
 * persistent launch then copies each probed (query, slice) into LDS and
 * probes the slice's region from it (the probe of mgv_hashjoin_probe_lds),
 * another inserts each built slice's region in LDS (the build of
 * mgv_hashjoin_build_lds), so neither fetches random lines; probe rows past
 * a full region are probed directly, build rows past one (and rows whose
 * chain leaves their slice) are inserted into the table afterwards; the
 * other queries run as mgv_batch_run.  A batch's builds and the probes that
 * read their tables go to two calls, builds first.
 * mgv_batch_plan_lds (host): validates as mgv_batch_plan; fills items
 * (capacity cap; *plan tells the count, also when items is NULL): scatter
 * tiles [0, nscatter), then one item per probed (query, slice) (.lo =
 * slice), then one per built (query, slice), then the other queries'
 * vectors of vector_rows rows (those queries are the small ones: short
 * vectors, such as 1024 rows, spread them over more waves); regions (4 words
 * per query: first region row, rows per region, slices F (0: not through
 * LDS), buckets per slice).
 * mgv_batch_run_lds: device copies of queries, items and regions; workspace
 * of plan->workspace_bytes (16-byte aligned), cleared by a fill kernel, so
 * the call can be captured into a graph.  launch.block / compact apply to
 * the other queries' launch.  No reference counterpart (the reference probes
 * each vector against the whole table, src/operators/hashjoin.cu:135-182). */
typedef struct mgv_lds_batch_plan {
    uint64_t nitems;             /* all items */
    uint64_t nscatter;           /* scatter tiles: items [0, nscatter) */
    uint64_t nparts;             /* probe (query, slice) items: [nscatter, nscatter + nparts) */
    uint64_t nbuild_parts;       /* build (query, slice) items: the next nbuild_parts */
    uint64_t build_rows;         /* rows of the LDS-built queries (their overflow list) */
    uint64_t region_rows;        /* rows of all regions */
    uint64_t workspace_bytes;
    uint32_t nqueries;
    uint32_t reserved;
} mgv_lds_batch_plan;
int mgv_batch_plan_lds(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, uint64_t min_rows,
                       mgv_work_item *items, uint64_t cap, uint64_t *regions, mgv_lds_batch_plan *plan);
int mgv_batch_run_lds(const mgv_query *queries, const mgv_work_item *items, const uint64_t *regions,
                      const mgv_lds_batch_plan *plan, void *ws, uint64_t ws_bytes, mgv_launch launch, mgv_stream_t s);

/* ------------------------------------------------------------ partitioned probe */

/* Workspace bytes of a partitioned probe of n rows with the given fan-out
 * (power of two in [2,512]) and partitioned columns (2 = key+val, 3 = with
 * row ids for the digest). */
int mgv_partition_workspace_size(uint64_t n, uint32_t fanout, uint32_t ncols, uint64_t *bytes);

/* Same results as mgv_hashjoin_probe, computed by first scattering the
 * (selected) probe rows into `fanout` partitions by the hash bits that pick
 * the bucket, then probing partition by partition with each XCD draining its
 * own partitions, so the table slice being probed stays in that XCD's L2.
 * The partitioning step is the counterpart of split::consume_warp
 * (src/operators/split.cu:54-118); the probe replaces
 * src/operators/hashjoin.cu:135-182.  Needs n < 2^31.  The workspace comes
 * from the caller; the call allocates nothing. */
int mgv_hashjoin_probe_partitioned(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, int64_t row_base,
                                   uint32_t flags, uint32_t fanout, void *workspace, uint64_t workspace_bytes,
                                   mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* Partitioned form of mgv_hashjoin_probe_groupby. */
int mgv_hashjoin_probe_groupby_partitioned(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred,
                                           int32_t bound, const int32_t *keys, const int32_t *vals, uint64_t n,
                                           uint64_t *gcnt, int64_t *gsum, uint32_t ngroups, uint32_t fanout,
                                           void *workspace, uint64_t workspace_bytes, mgv_launch launch,
                                           mgv_stream_t s);

/* Diagnostic: out8[x] += number of `grid` one-wave blocks that ran on XCD x. */
int mgv_debug_xcc_histogram(uint32_t *out8, uint32_t grid, mgv_stream_t s);

/* Build through LDS-resident table slices (same table contents, as sets
 * of {key, payload} per chain, as mgv_hashjoin_build): the rows (optionally
 * only those with pred <= bound) are partitioned by table slice as in
 * mgv_hashjoin_probe_lds, then one block per slice copies the slice into
 * LDS, inserts the slice's rows with LDS compare-and-swap (each bucket's
 * slots claimed in order), and writes it back; rows whose chain runs past
 * their slice, and rows whose partition region was full (duplicate or
 * skewed keys), are inserted into the global table afterwards.  Existing
 * contents of the table are kept.  No other build may write the same table
 * concurrently.  MGV_PAYLOAD_ROWID runs mgv_hashjoin_build.  The
 * partitionings are count-free by default (as mgv_hashjoin_probe_lds);
 * MGV_LAUNCH_COUNTED_PART selects the counted ones (same table contents,
 * at most 2048 slices).  Same bucket and fan-out limits as
 * mgv_hashjoin_probe_lds; n < 2^31; the workspace
 * (mgv_build_lds_workspace_size, enough for either partitioning) holds both
 * partitionings and the overflow list (about 40 bytes per row).  Replaces
 * hashjoin_builder::consume_warp (src/operators/hashjoin.cu:41-60). */
int mgv_build_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes);
int mgv_hashjoin_build_lds(uint64_t *slots, uint32_t nbuckets, const int32_t *keys, const int32_t *payload,
                           int payload_mode, const int32_t *pred, int32_t bound, uint64_t n, int64_t row_base,
                           uint32_t fanout, void *workspace, uint64_t workspace_bytes, mgv_result *res,
                           mgv_launch launch, mgv_stream_t s);

/* Probe + aggregate through LDS-resident table slices (same results as
 * mgv_hashjoin_probe without flags): the probe rows (optionally only those
 * with pred <= bound) are hash-partitioned twice -- `fanout` partitions,
 * then nbuckets / S partitions whose table slices of S <= 2048 buckets
 * (128 KiB) fit in LDS -- and each slice is probed from LDS by one block.
 * The partitionings need no count passes: each partition is a region of
 * fixed capacity (a few standard deviations above its expected size)
 * filled by one atomic reservation per tile, and rows beyond a full region
 * (skewed keys) go to an overflow area that is probed from the table
 * directly.  MGV_LAUNCH_COUNTED_PART selects exact counted partitionings
 * instead (same results).
 * nbuckets must split into power-of-two slices of at most 2048 buckets:
 * when nbuckets / fanout is such a slice there is one partitioning level
 * only, otherwise the second level cuts each of the fanout partitions into
 * a power of two of at most 1024 slices (at most 32768 in all; 2048 with
 * MGV_LAUNCH_COUNTED_PART).  n < 2^31.  The workspace
 * (size from mgv_probe_lds_workspace_size) holds both partitionings (about
 * 32 bytes per row with their overflow areas).  The table must come from
 * mgv_hashjoin_build or a batch build (mgv_batch_run), which fill each
 * bucket's slots in order: the LDS copy of a slice keeps per bucket only
 * the keys, the payloads and the number of filled slots.  For a table of
 * 2^B >= 2^18 buckets the copy holds per slot a 16-bit tag instead of the
 * key: the slot's chain offset (B - 16 bits) and the low 32 - B bits of its
 * key's hash, which together with the bucket identify the key exactly
 * (chains longer than the offset field continue in global memory).  Replaces
 * hashjoin::consume_warp (src/operators/hashjoin.cu:135-182) fused with
 * aggregation (src/operators/aggregation.cu:27-66). */
int mgv_probe_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes);
int mgv_hashjoin_probe_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                           const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout, void *ws,
                           uint64_t ws_bytes, mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* Probe + group-by through LDS-resident table slices (same results as
 * mgv_hashjoin_probe_groupby): the probe rows (optionally only those with
 * pred <= bound) are partitioned as in mgv_hashjoin_probe_lds, each slice
 * of at most 1024 buckets is probed from LDS by one block that aggregates
 * COUNT and SUM(vals) per group (group = payload of the matching slot,
 * groups >= ngroups dropped) in an LDS table flushed into gcnt / gsum with
 * atomics; rows whose partition overflowed (skewed keys) are probed from the
 * table directly with the same per-block aggregation.  ngroups <= 4096;
 * nbuckets must split into power-of-two slices of at most 1024 buckets;
 * workspace from mgv_probe_groupby_lds_workspace_size.  No reference
 * counterpart for the group-by; the probe replaces
 * src/operators/hashjoin.cu:135-182. */
int mgv_probe_groupby_lds_workspace_size(uint64_t n, uint32_t nbuckets, uint32_t fanout, uint64_t *bytes);
int mgv_hashjoin_probe_groupby_lds(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                                   uint32_t ngroups, uint32_t fanout, void *ws, uint64_t ws_bytes, mgv_launch launch,
                                   mgv_stream_t s);

/* mgv_hashjoin_probe_lds without a workspace argument and without any
 * allocation: the partitionings use a scratch array of
 * MGV_PROBE_NOALLOC_SCRATCH_BYTES that the library holds in each GPU's
 * memory from mgv_init until the process ends.  Calls whose workspace
 * (mgv_probe_lds_workspace_size) exceeds it fail with MGV_E_CAPACITY before
 * any launch.  Every call on a GPU uses that GPU's one array: calls that
 * overlap on one GPU (different streams, or host threads) race and must be
 * serialised by the caller.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under stream capture, serving
 * loops with a fixed footprint). */
#define MGV_PROBE_NOALLOC_SCRATCH_BYTES (64ull << 20)
int mgv_hashjoin_probe_lds_noalloc(const uint64_t *slots, uint32_t nbuckets, const int32_t *pred, int32_t bound,
                                   const int32_t *keys, const int32_t *vals, uint64_t n, uint32_t fanout,
                                   mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* ------------------------------------------------------------ select / aggregate */

/* SELECT COUNT(*), SUM(val) WHERE pred <= bound.  Replaces the
 * unstable_select -> aggregation chain (src/operators/select3.cu:19-85,
 * src/operators/aggregation.cu:19-66, plan src/main9.cu:202-226).  val may
 * equal pred (the reference's single-column form). */
int mgv_select_sum(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, mgv_result *res,
                   mgv_launch launch, mgv_stream_t s);

/* Stream compaction: out = {val[i] : pred[i] <= bound} in unspecified order
 * (res->count = selected rows, res->overflow = rows beyond cap).  Replaces
 * unstable_select with its gpu_to_cpu parent (src/operators/select3.cu:19-270,
 * src/buffer.cuh:92-134, plan src/main9.cu:150-200). */
int mgv_select_materialize(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound, int32_t *out,
                           uint64_t cap, mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* GROUP BY key: cnt[key] += 1, sum[key] += val, keys in [0, ngroups)
 * (other keys counted in res->errors when res != NULL).  No reference
 * counterpart (aggregation.cu is scalar only). */
int mgv_groupby(const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                uint32_t ngroups, mgv_result *res, mgv_launch launch, mgv_stream_t s);

/* GROUP BY over many groups (more than fit an LDS table): the rows are
 * scattered into key ranges of 8192 groups (range = key >> 13), then each
 * range is aggregated in an LDS table by persistent blocks that flush one
 * pair of atomics per touched group.  Same results and error counting as
 * mgv_groupby (which it calls for ngroups <= 4096); n < 2^31; ngroups <=
 * 2^22.  The workspace (device, size from mgv_groupby_workspace_size) holds
 * the partitioned copy of keys and vals (8 bytes per row).  No reference
 * counterpart (aggregation.cu is scalar only). */
int mgv_groupby_workspace_size(uint64_t n, uint32_t ngroups, uint64_t *bytes);
int mgv_groupby_partitioned(const int32_t *keys, const int32_t *vals, uint64_t n, uint64_t *gcnt, int64_t *gsum,
                            uint32_t ngroups, void *ws, uint64_t ws_bytes, mgv_result *res, mgv_launch launch,
                            mgv_stream_t s);

/* ------------------------------------------------------------ split */

/* The split operator (src/operators/split.cu:54-118, partition function
 * hash3(x) = x & 1 at split.cu:49-51): every row of a vector whose first
 * column x has (x & 1) == p, and which passes the optional filter
 * pred <= bound (pred NULL: every row), is appended to parent p's output
 * columns.  Rows whose destination has no parent are dropped, as in the
 * reference: with nparents == 1 the odd rows, and parents 2.. receive
 * nothing.
 *   cols      host array of ncols device columns of n rows (the row's
 *             values; column 0 decides the destination)
 *   out       host array of nparents * ncols device columns; parent p's
 *             column c is out[p * ncols + c], `cap` rows each
 *   counts    device array of nparents append cursors (rows so far, 8-byte
 *             aligned): a row
 *             for parent p goes to position counts[p]++ of its columns;
 *             rows at positions >= cap are not written but still counted,
 *             so counts[p] > cap after the call reports an overflow
 * The order of the rows inside a parent's output is unspecified.  The
 * launch's grid (0 = one persistent grid) is honoured, its block is not
 * (one column: blocks of 512 threads take tiles of 16384 rows; two to four
 * columns: 1024 threads, 8192 rows). */
#define MGV_SPLIT_MAX_COLS 4
#define MGV_SPLIT_MAX_PARENTS 8
int mgv_split(const int32_t *pred, int32_t bound, const int32_t *const *cols, uint32_t ncols, uint64_t n,
              uint32_t nparents, int32_t *const *out, uint64_t cap, uint64_t *counts, mgv_launch launch,
              mgv_stream_t s);

#ifdef __cplusplus
}
#endif

#endif /* MGV_H_ */
