/* mgv_ops.h -- C ABI of the C++ operator layer (libmgv_ops.so).
 *
 * These entry points assemble the reference's query plans out of the C++
 * drop-in operators (d_operator_t / h_operator_t / generator / exchange /
 * gpu_to_cpu / materializer, multigpu-vector_amd/host/include/mgv/) and run
 * them, so the operator lifecycle can be driven from any FFI.  Each plan is
 * the one-GPU form of a plan builder in src/main9.cu; columns may live in
 * device memory (device = 0) or in host memory (device = -1, copied to the
 * GPU vector by vector through pinned double buffers).
 *
 * Every function returns 0 (MGV_OK) or an error code of mgv.h; the message
 * of the last failure on the calling thread is mgv_ops_last_error().
 */
#ifndef MGV_OPS_H
#define MGV_OPS_H

#include "mgv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* buffer_manager<int32_t>::init(size, buff_buffer_size, keep_threshold)
 * (src/buffer_manager.cu:38-211): nbuffers device buffers and nbuffers
 * pinned host buffers of 4 Mi rows each. */
int mgv_ops_init(uint32_t nbuffers, uint32_t buff_buffer_size, uint32_t keep_threshold);
/* buffer_manager<int32_t>::destroy (src/buffer_manager.cu:213). */
int mgv_ops_shutdown(void);
const char *mgv_ops_last_error(void);
/* Buffers of the manager's device pool: free in the pool (its stack and the
 * buffers handed back from the host), held in the host-side cache of
 * h_get_buffer, and the pool's size; and free pinned host buffers.  With no
 * plan running, device_free + device_cached == device_total. */
int mgv_ops_pool_stats(uint32_t *device_free, uint32_t *device_cached, uint32_t *device_total, uint32_t *host_free);

/* generator -> exchange -> unstable_select(less_eq_than(thres)) ->
 * aggregation -> gpu_to_cpu -> exchange -> materializer
 * (create_sum_select_pipeline, src/main9.cu:202-226).  The vectors carry
 * {pred, val} (val may equal pred: the reference's single-column form).
 * *sum = exact int64 SUM(val WHERE pred <= thres); dst[0] (if dst) = its
 * low 32 bits, the value the reference's int32 aggregation delivers. */
int mgv_ops_sum_select(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                       int device, uint32_t grid, uint32_t block, int32_t *dst, int64_t *sum);

/* generator -> exchange -> unstable_select(less_eq_than(thres)) ->
 * gpu_to_cpu -> exchange -> materializer(dst)
 * (create_select_pipeline, src/main9.cu:150-174); with apply_select = 0 the
 * select is left out (create_select_all_pipeline, src/main9.cu:176-200).
 * dst (host, capacity dst_cap >= n, may be NULL) receives the selected rows
 * of val in unspecified order; *count = selected rows. */
int mgv_ops_select(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, int apply_select,
                   uint32_t vector_rows, int device, uint32_t grid, uint32_t block, int32_t *dst, uint64_t dst_cap,
                   uint64_t *count);

/* generator -> exchange -> unstable_select(less_eq_than(thres)) ->
 * split(x & 1) -> {union_all, union_all} (one union_all, both parents of the
 * split) -> gpu_to_cpu -> exchange -> materializer(dst)
 * (create_select_pipeline_exp, src/main9.cu:228-258, on one GPU).  The split
 * routes each selected row by its parity and the union_all brings both
 * halves back together, so the contract is mgv_ops_select's: dst (if not
 * NULL, capacity dst_cap >= n) receives the selected rows of val in
 * unspecified order, *count = selected rows. */
int mgv_ops_select_split(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                         int device, uint32_t grid, uint32_t block, int32_t *dst, uint64_t dst_cap, uint64_t *count);

/* generator -> exchange -> unstable_select(less_eq_than(thres)) ->
 * split(x & 1) -> {aggregation -> gpu_to_cpu -> exchange -> materializer} x 2:
 * a split with two different parents (the shape of hash_exchange,
 * src/main9.cu:104-148, with both sides on one GPU).  sums[p] = exact int64
 * SUM(val WHERE pred <= thres AND (val & 1) == p), counts[p] = those rows;
 * sums and counts are host arrays of 2. */
int mgv_ops_split_sums(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                       int device, uint32_t grid, uint32_t block, int64_t *sums, uint64_t *counts);

/* Multi-stream fan-in: generator -> exchange with nstreams consumers (each
 * with its own streams, vectors going to whichever is free) -> one
 * unstable_select(less_eq_than(thres)) per consumer -> union_all(nstreams)
 * -> [aggregation ->] gpu_to_cpu -> exchange -> materializer (the
 * union_all fan-in of union_all.cu:36-57, src/main9.cu:104-148, on one
 * GPU).  nstreams in [1, 16].  _sum: *sum = exact int64 SUM(val WHERE pred
 * <= thres); _select: dst (if not NULL, capacity dst_cap >= n) receives the
 * selected rows of val in unspecified order, *count = selected rows. */
int mgv_ops_sum_select_streams(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres,
                               uint32_t vector_rows, int device, uint32_t grid, uint32_t block, uint32_t nstreams,
                               int64_t *sum);
int mgv_ops_select_streams(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres, uint32_t vector_rows,
                           int device, uint32_t grid, uint32_t block, uint32_t nstreams, int32_t *dst,
                           uint64_t dst_cap, uint64_t *count);

/* create_select_pipeline (aggregate = 0) / create_sum_select_pipeline
 * (aggregate = 1) with npipelines launch confs on one GPU
 * (src/main9.cu:150-174, 202-226): generator -> exchange with one consumer
 * per conf -> unstable_select(less_eq_than(thres)) -> [aggregation ->]
 * gpu_to_cpu -> mem_move -> union_all_cpu(npipelines) -> exchange ->
 * materializer(dst).  npipelines in [1, 16].  aggregate = 0: dst (if not
 * NULL, dst_cap >= n) receives the selected rows of val, *value = their
 * number; aggregate = 1: *value = exact int64 SUM(val WHERE pred <= thres),
 * dst (if not NULL, dst_cap >= npipelines) the low 32 bits of each
 * pipeline's partial sum, as the reference's materializer stores them. */
int mgv_ops_select_pipelines(const int32_t *pred, const int32_t *val, uint64_t n, int32_t thres,
                             uint32_t vector_rows, int device, uint32_t grid, uint32_t block, uint32_t npipelines,
                             int aggregate, int32_t *dst, uint64_t dst_cap, int64_t *value);

/* Build: generator(R) -> exchange -> [select(pred_r <= bound) ->]
 * hashjoin_builder; probe: generator(L) -> exchange -> [select(pred_l <=
 * bound) ->] hashjoin -> gpu_to_cpu -> exchange -> materializer
 * (create_count_join_1G_pipeline, src/main9.cu:326-367, and its select
 * variant 260-324 on one GPU).  pred_r / pred_l may be NULL (no select).
 * *count = join matches (int64); dst[0] (if dst) = its low 32 bits. */
int mgv_ops_count_join(const int32_t *R, const int32_t *pred_r, uint64_t rn, const int32_t *L, const int32_t *pred_l,
                       uint64_t ln, int32_t bound, uint32_t vector_rows, int device, uint32_t grid, uint32_t block,
                       int32_t *dst, uint64_t *count);

/* The plain per-vector path for a batch: for each query, for each vector of
 * vector_rows rows, one launch of the query's operator chain
 * (launch_consume_pipeline, src/operators/exchange.cu:160-194) on stream s.
 * `queries` is a host array (device pointers inside), same contract as
 * mgv_batch_run; results accumulate in each query's res / gcnt / gsum.
 * Builds run before probes when the queries list them first. */
int mgv_ops_run_plain(const mgv_query *queries, uint32_t nqueries, uint32_t vector_rows, mgv_launch launch,
                      mgv_stream_t s);

/* ------------------------------------------------------------ serving loop */

/* A request of the serving loop over resident columns (device pointers). */
#define MGV_SERVE_JOIN_SUM 1   /* build S(bkey, bval) then COUNT, SUM(val + w) of R JOIN S WHERE pred <= bound */
#define MGV_SERVE_GROUPBY 2    /* GROUP BY key (in [0, ngroups)): COUNT, SUM(val) */

typedef struct mgv_serve_request {
    uint32_t kind;
    int32_t bound;
    const int32_t *pred;       /* JOIN_SUM: filter column (NULL = no filter) */
    const int32_t *key;        /* probe keys / group keys */
    const int32_t *val;
    uint64_t n;                /* probe / group-by rows */
    const int32_t *bkey;       /* JOIN_SUM: build keys (unique) and payloads */
    const int32_t *bval;
    uint64_t bn;
    uint32_t ngroups;          /* GROUPBY */
    uint32_t reserved;
    double arrival_s;          /* arrival time, seconds after the start of the run */
} mgv_serve_request;

typedef struct mgv_serve_result {
    uint64_t count;            /* join matches / grouped rows */
    int64_t sum;
    uint64_t digest;           /* GROUPBY: digest of the per-group (COUNT, SUM) arrays; JOIN_SUM: 0 */
    uint64_t footprint;        /* HBM bytes the query held while it ran */
    double admit_s;            /* admitted (memory reserved, launched) */
    double done_s;             /* result on the host */
    double latency_s;          /* done_s - arrival_s */
    uint32_t batch;            /* index of the batch that ran it */
    uint32_t errors;           /* build rejects / out-of-range group keys */
} mgv_serve_result;

typedef struct mgv_serve_config {
    uint64_t budget_bytes;     /* HBM budget for hash tables + group arrays */
    uint32_t max_inflight;     /* batches on the GPU at once (one stream each), 1..8 */
    uint32_t max_batch;        /* queries per batch, 1..1024 */
    uint32_t vector_rows;      /* largest work item of the fused kernel (a multiple of 256) */
    uint32_t reserved;
    double load_factor;        /* hash tables */
    mgv_launch launch;
} mgv_serve_config;

typedef struct mgv_serve_stats {
    double elapsed_s;          /* first arrival to last result */
    double queries_per_s;
    double p50_s, p99_s, max_s, mean_s;
    uint32_t batches;
    uint32_t max_inflight_seen;
    uint64_t peak_bytes;       /* peak reserved HBM of the budget */
    double mean_batch;         /* queries per batch */
    uint64_t ring_results;     /* results that reached the host through the completion ring */
} mgv_serve_stats;

/* Continuous-batching request loop.  Requests (sorted by arrival_s) are
 * released at their arrival time; a waiting request is admitted when its
 * footprint (hash table at load_factor, or group arrays, + result) fits the
 * free HBM budget and a stream slot is free; the admitted requests (while
 * their probe, group-by and build rows add up to at most 2^26; a larger
 * request runs alone) form one batch, each request cut into about
 * sqrt(rows / 2) work items (at most 8192, of at most vector_rows rows):
 * one persistent launch for its builds and one for its probes and
 * group-bys (the fused kernel of mgv_batch_run; a group-by over more than
 * 4096 groups runs as mgv_groupby_partitioned, its workspace counted in its
 * footprint), after which the results are copied to pinned host memory
 * and each request's {id, result} is published on a device-written
 * completion ring (mgv_ring, the gpu_to_cpu ring of src/operators/
 * gpu_to_cpu.cu:42-58,149-165), which the loop polls.  Up to max_inflight
 * batches overlap on the GPU; a batch's memory is returned when all its
 * requests have come off the ring.
 * Admission is FIFO (a request that does not fit blocks the ones behind it).
 * Fails with MGV_E_CAPACITY for a request larger than the whole budget. */
int mgv_serve_run(const mgv_serve_request *requests, uint32_t n, const mgv_serve_config *config,
                  mgv_serve_result *results, mgv_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* MGV_OPS_H */
