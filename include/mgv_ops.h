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

#ifdef __cplusplus
}
#endif

#endif /* MGV_OPS_H */
