/*
 * mgv_oracle.h -- CPU oracle for the select -> hash-join -> aggregate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * operator semantics (chapeiro/multigpu-vector) and of the synthetic-data
 * generator.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline.  The
 * product path (libmgv.so) never links, loads or calls it.
 *
 * Pinning status of each function (what makes a plausible mistake fail):
 *   orc_philox4x32_10   Random123 known-answer vectors; rocRAND's
 *                       philox4x32_10_engine on sampled counters.
 *   orc_fmix32          sklearn.utils.murmurhash3_32 (full MurmurHash3_x86_32,
 *                       whose last step is fmix32); bijectivity on samples.
 *   orc_feistel_perm    bijection check (sorted output == 0..n-1), numpy
 *                       re-implementation in tests/golden.
 *   orc_zipf_cdf        closed form for s=1 (harmonic numbers), numpy.
 *   orc_select_*        brute-force numpy count/sum; int32 wrap == (int32)int64.
 *   orc_hashjoin_*      closed form matches = #{r : R.k < n_S} for permutation
 *                       builds; nested-loop brute force at <= 2^12 rows;
 *                       self-join sum_v hist[v]^2 for duplicate keys.
 *   orc_groupby_*       numpy bincount; sum over groups == total.
 *   orc_split           partition invariants: the parents' rows are a
 *                       partition of the kept rows, parent p holds exactly
 *                       the rows with x % 2 == p (Python modulo), one parent
 *                       keeps only the even rows, parents >= 2 stay empty.
 * No function is "parity unpinned".
 */
#ifndef MGV_ORACLE_H_
#define MGV_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- datagen */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  out = bijection(ctr; key). */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Seed of (config, column): 0x4D47560000000000 | config<<8 | column. */
uint64_t orc_seed(uint32_t config, uint32_t column);

/* 32-bit word i of the stream (seed, stream): word (i&3) of
 * philox(ctr = {lo(i>>2), hi(i>>2), stream, 0}, key = {lo(seed), hi(seed)}). */
uint32_t orc_word(uint64_t seed, uint32_t stream, uint64_t i);

/* dst[i] = word(i) >> (32-bits), bits in [0,32]. */
void     orc_gen_bits(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t bits);

/* dst[i] = (word(i) * range) >> 32 -- uniform in [0, range). */
void     orc_gen_range(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t range);

/* Keyed balanced Feistel bijection on [0, 2^nbits), nbits even, 4 rounds. */
uint32_t orc_feistel(uint32_t x, uint32_t nbits, uint64_t seed, uint32_t stream);

/* Bijection on [0, range) by cycle-walking orc_feistel over the smallest even
 * width >= log2(range) (and >= 2). */
uint32_t orc_perm(uint32_t x, uint32_t range, uint64_t seed, uint32_t stream);

/* dst[i] = orc_perm(i, range) for i < n (n <= range). */
void     orc_gen_perm(int32_t *dst, uint64_t n, uint32_t range, uint64_t seed, uint32_t stream);

/* Zipf(s) thresholds over n ranks: T[r] = floor(2^32 * F(r)), F(r) = sum_{j<=r}
 * (j+1)^-s / H, summed in ascending order in double; T[n-1] = 2^32. */
void     orc_zipf_cdf(uint64_t *T, uint32_t n, double s);

/* rank = smallest r with u < T[r]. */
uint32_t orc_zipf_rank(uint32_t u, const uint64_t *T, uint32_t n);

/* dst[i] = map(rank(word(i))) where map is orc_perm(., n, pseed, pstream) if
 * pseed != 0 else the identity. */
void     orc_gen_zipf(int32_t *dst, uint64_t cnt, uint64_t seed, uint32_t stream,
                      const uint64_t *T, uint32_t n, uint64_t pseed, uint32_t pstream);

/* ---------------------------------------------------------------- hash */

/* MurmurHash3 fmix32 finaliser, as hashMurmur (hashjoin.cu:4-11). */
uint32_t orc_fmix32(uint32_t x);

/* 64-bit mixer (splitmix64 finaliser) for order-independent pair digests. */
uint64_t orc_mix64(uint64_t x);

/* ---------------------------------------------------------------- select */

/* stable_select_cpu (main9.cu:48-52) with less_eq_than (select3.cuh:10-23):
 * copies x with x <= bound into out (if out) and returns the count. */
uint64_t orc_select(const int32_t *x, uint64_t n, int32_t bound, int32_t *out);

/* sum_select_cpu generalised to a payload column (main9.cu:56-61):
 * count = #{i : pred[i] <= bound}; sum64 = sum val[i] over them (int64);
 * sum32 = the same sum accumulated in int32 with wrap-around, exactly as the
 * reference's int32 accumulator (aggregation.cuh:12, main9.cu:57). */
void     orc_select_sum(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound,
                        uint64_t *count, int64_t *sum64, int32_t *sum32);

/* Same as orc_select_sum, using nthreads host threads (per-thread partials). */
void     orc_select_sum_mt(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound,
                           uint64_t *count, int64_t *sum64, int nthreads);

/* ---------------------------------------------------------------- join */

/* Chained hash table of sum_hashjoin_select_cpu (main8.cu:84-103) and of
 * hashjoin_builder/hashjoin (hashjoin.cu:41-60,135-182): heads[2^log_buckets]
 * initialised to -1, next[] links, bucket = fmix32(key) & (2^log_buckets - 1),
 * insertion pushes at the head of the chain. */
typedef struct orc_chained_table {
    int32_t  *heads;
    int32_t  *next;
    const int32_t *keys;
    const int32_t *payload;   /* may be NULL */
    uint32_t  log_buckets;
    uint64_t  n;
} orc_chained_table;

int      orc_ct_build(orc_chained_table *t, const int32_t *keys, const int32_t *payload,
                      uint64_t n, uint32_t log_buckets);
/* The same table built by nthreads threads with an atomic exchange on the
 * bucket head (the reference's device build, hashjoin.cu:54-55); chain
 * order differs from orc_ct_build, chain contents do not. */
int      orc_ct_build_mt(orc_chained_table *t, const int32_t *keys, const int32_t *payload,
                         uint64_t n, uint32_t log_buckets, int nthreads);
void     orc_ct_free(orc_chained_table *t);

/* Join against a build side with distinct keys in [0, kmax) by direct
 * addressing: COUNT and int64 SUM(rv[r] + sw[s]) over the matching rows
 * (optionally pred[r] <= bound).  For C2 (build keys a permutation of
 * [0, 2^24)) this is the closed form COUNT = #{r : R.k < 2^24}, SUM = sum
 * of R.v + w[perm^-1(R.k)].  Returns -1 for an out-of-range or repeated
 * build key.  sw, pred and rv may be NULL. */
int      orc_join_dense(const int32_t *sk, const int32_t *sw, uint64_t ns, uint32_t kmax,
                        const int32_t *pred, int32_t bound, const int32_t *rk, const int32_t *rv,
                        uint64_t nr, uint64_t *count, int64_t *sum, int nthreads);

/* Probe: for every probe row r (optionally filtered by pred[r] <= bound when
 * pred != NULL), for every build row s with key[s] == rk[r]:
 *   count += 1; sum += (int64)rv[r] + (int64)payload[s]  (rv/payload may be NULL -> 0)
 *   digest_sum += mix64(r << 32 | s); digest_xor ^= mix64(r << 32 | s)
 * count32 is the reference's int32 COUNT (hashjoin.cuh:41 res), wrapped. */
typedef struct orc_join_result {
    uint64_t count;
    int64_t  sum;
    uint64_t digest_sum;
    uint64_t digest_xor;
    int32_t  count32;
} orc_join_result;

void     orc_ct_probe(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                      const int32_t *rk, const int32_t *rv, uint64_t nr, uint64_t r_base,
                      orc_join_result *res);
void     orc_ct_probe_mt(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                         const int32_t *rk, const int32_t *rv, uint64_t nr,
                         orc_join_result *res, int nthreads);

/* Group-by after a filtered join: for every match, group g = payload[s]
 * (in [0, ngroups)), cnt[g] += 1, sum[g] += rv[r].  cnt/sum are zeroed here. */
void     orc_ct_probe_groupby(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                              const int32_t *rk, const int32_t *rv, uint64_t nr,
                              uint64_t *cnt, int64_t *sum, uint32_t ngroups);
/* orc_ct_probe_groupby on nthreads threads (per-thread group arrays). */
int      orc_ct_probe_groupby_mt(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                                 const int32_t *rk, const int32_t *rv, uint64_t nr,
                                 uint64_t *cnt, int64_t *sum, uint32_t ngroups, int nthreads);

/* Plain group-by: cnt[key[i]] += 1, sum[key[i]] += val[i]; keys in [0, ngroups). */
void     orc_groupby(const int32_t *key, const int32_t *val, uint64_t n,
                     uint64_t *cnt, int64_t *sum, uint32_t ngroups);

/* Self-join count, sum_selfjoin_select_cpu (main9.cu:63-69): sum over values v
 * with v <= bound of hist[v]^2, for values in [0, vmax].  Returns int64; the
 * reference's int32 result is its low 32 bits. */
/* split (src/operators/split.cu:54-118; hash3(x) = x & 1 at 49-51): row i
 * (kept when pred is NULL or pred[i] <= bound) goes to parent cols[0][i] & 1
 * when that is < nparents, else it is dropped.  Parent p's rows are written
 * in input order to out[p * ncols + c] (c < ncols); counts[p] = its rows.
 * The outputs must hold n rows each. */
void     orc_split(const int32_t *pred, int32_t bound, const int32_t *const *cols, uint32_t ncols, uint64_t n,
                   uint32_t nparents, int32_t *const *out, uint64_t *counts);

int64_t  orc_selfjoin_count(const int32_t *x, uint64_t n, int32_t bound, uint32_t vmax);

#ifdef __cplusplus
}
#endif

#endif
