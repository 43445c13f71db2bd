/*
 * mgv_oracle.c -- CPU oracle (test infrastructure; see mgv_oracle.h).
 *
 * Every function names the reference lines whose semantics it restates.
 * Reference paths are relative to the chapeiro/multigpu-vector tree.
 */
#include "mgv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ Philox */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t orc_seed(uint32_t config, uint32_t column) {
    return 0x4D47560000000000ull | ((uint64_t)config << 8) | (uint64_t)column;
}

uint32_t orc_word(uint64_t seed, uint32_t stream, uint64_t i) {
    uint64_t blk = i >> 2;
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[i & 3];
}

/* Generate words for [0, n) block by block (one Philox call per 4 words). */
static void gen_words(uint64_t n, uint64_t seed, uint32_t stream,
                      void (*emit)(void *ctx, uint64_t i, uint32_t w), void *ctx) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint64_t blk = 0; blk * 4 < n; ++blk) {
        uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
        uint32_t out[4];
        orc_philox4x32_10(ctr, key, out);
        for (uint32_t j = 0; j < 4 && blk * 4 + j < n; ++j) emit(ctx, blk * 4 + j, out[j]);
    }
}

typedef struct { int32_t *dst; uint32_t bits; uint32_t range; } gen_ctx;

static void emit_bits(void *c, uint64_t i, uint32_t w) {
    gen_ctx *g = (gen_ctx *)c;
    g->dst[i] = g->bits ? (int32_t)(w >> (32 - g->bits)) : 0;
}

static void emit_range(void *c, uint64_t i, uint32_t w) {
    gen_ctx *g = (gen_ctx *)c;
    g->dst[i] = (int32_t)(uint32_t)(((uint64_t)w * g->range) >> 32);
}

void orc_gen_bits(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t bits) {
    gen_ctx g = {dst, bits, 0};
    gen_words(n, seed, stream, emit_bits, &g);
}

void orc_gen_range(int32_t *dst, uint64_t n, uint64_t seed, uint32_t stream, uint32_t range) {
    gen_ctx g = {dst, 0, range};
    gen_words(n, seed, stream, emit_range, &g);
}

/* ------------------------------------------------------------------ hashes */

/* hashMurmur, hashjoin.cu:4-11 (MurmurHash3 fmix32, A. Appleby). */
uint32_t orc_fmix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

/* splitmix64 finaliser (S. Vigna). */
uint64_t orc_mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

/* ------------------------------------------------------------------ Feistel */

static uint32_t feistel_round_key(uint64_t seed, uint32_t stream, uint32_t r) {
    uint32_t inner = (uint32_t)(seed >> 32) + stream * 0x85EBCA6Bu + r * 0x9E3779B9u;
    return orc_fmix32((uint32_t)seed ^ orc_fmix32(inner));
}

uint32_t orc_feistel(uint32_t x, uint32_t nbits, uint64_t seed, uint32_t stream) {
    uint32_t h = nbits / 2;
    uint32_t m = (h >= 32) ? 0xffffffffu : ((1u << h) - 1u);
    uint32_t L = x >> h, R = x & m;
    for (uint32_t r = 0; r < 4; ++r) {
        uint32_t F = orc_fmix32(R ^ feistel_round_key(seed, stream, r)) & m;
        uint32_t nL = R;
        R = L ^ F;
        L = nL;
    }
    return (L << h) | R;
}

static uint32_t perm_width(uint32_t range) {
    uint32_t nb = 2;
    while (nb < 32 && ((uint64_t)1 << nb) < range) nb += 2;
    return nb;
}

uint32_t orc_perm(uint32_t x, uint32_t range, uint64_t seed, uint32_t stream) {
    uint32_t nb = perm_width(range);
    uint32_t y = x;
    do { y = orc_feistel(y, nb, seed, stream); } while (y >= range);
    return y;
}

void orc_gen_perm(int32_t *dst, uint64_t n, uint32_t range, uint64_t seed, uint32_t stream) {
    for (uint64_t i = 0; i < n; ++i) dst[i] = (int32_t)orc_perm((uint32_t)i, range, seed, stream);
}

/* ------------------------------------------------------------------ Zipf */

void orc_zipf_cdf(uint64_t *T, uint32_t n, double s) {
    double H = 0.0;
    for (uint32_t r = 0; r < n; ++r) H += (s == 1.0) ? 1.0 / (double)(r + 1) : 1.0 / pow((double)(r + 1), s);
    double acc = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
        acc += (s == 1.0) ? 1.0 / (double)(r + 1) : 1.0 / pow((double)(r + 1), s);
        double f = acc / H * 4294967296.0;
        uint64_t t = (f >= 4294967296.0) ? 4294967296ull : (uint64_t)f;
        T[r] = t;
    }
    if (n) T[n - 1] = 4294967296ull;
}

uint32_t orc_zipf_rank(uint32_t u, const uint64_t *T, uint32_t n) {
    uint32_t lo = 0, hi = n - 1;           /* answer in [lo, hi]; T[n-1] > u always */
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)u < T[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

void orc_gen_zipf(int32_t *dst, uint64_t cnt, uint64_t seed, uint32_t stream,
                  const uint64_t *T, uint32_t n, uint64_t pseed, uint32_t pstream) {
    for (uint64_t i = 0; i < cnt; ++i) {
        uint32_t r = orc_zipf_rank(orc_word(seed, stream, i), T, n);
        dst[i] = (int32_t)(pseed ? orc_perm(r, n, pseed, pstream) : r);
    }
}

/* ------------------------------------------------------------------ select */

/* stable_select_cpu, main9.cu:48-52; predicate less_eq_than, select3.cuh:10-23. */
uint64_t orc_select(const int32_t *x, uint64_t n, int32_t bound, int32_t *out) {
    uint64_t k = 0;
    for (uint64_t j = 0; j < n; ++j)
        if (x[j] <= bound) { if (out) out[k] = x[j]; ++k; }
    return k;
}

/* sum_select_cpu, main9.cu:56-61, with the summed column separated from the
 * predicate column (the reference sums the predicate column itself; a test
 * passes the same pointer twice to get exactly that). */
void orc_select_sum(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound,
                    uint64_t *count, int64_t *sum64, int32_t *sum32) {
    uint64_t c = 0;
    int64_t s = 0;
    uint32_t s32 = 0;                     /* int32 accumulation with wrap-around */
    for (uint64_t j = 0; j < n; ++j) {
        if (pred[j] <= bound) {
            ++c;
            s += val[j];
            s32 += (uint32_t)val[j];
        }
    }
    if (count) *count = c;
    if (sum64) *sum64 = s;
    if (sum32) *sum32 = (int32_t)s32;
}

typedef struct {
    const int32_t *pred, *val;
    uint64_t lo, hi;
    int32_t bound;
    uint64_t count;
    int64_t sum;
} sel_task;

static void *sel_worker(void *p) {
    sel_task *t = (sel_task *)p;
    uint64_t c = 0;
    int64_t s = 0;
    for (uint64_t j = t->lo; j < t->hi; ++j) {
        int q = t->pred[j] <= t->bound;
        c += (uint64_t)q;
        s += q ? t->val[j] : 0;
    }
    t->count = c;
    t->sum = s;
    return NULL;
}

void orc_select_sum_mt(const int32_t *pred, const int32_t *val, uint64_t n, int32_t bound,
                       uint64_t *count, int64_t *sum64, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    pthread_t th[256];
    sel_task tk[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        tk[t].pred = pred; tk[t].val = val; tk[t].bound = bound;
        tk[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
        tk[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, sel_worker, &tk[t]);
    }
    uint64_t c = 0;
    int64_t s = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        c += tk[t].count;
        s += tk[t].sum;
    }
    *count = c;
    *sum64 = s;
}

/* ------------------------------------------------------------------ join */

/* Build of sum_hashjoin_select_cpu (main8.cu:84-91) / hashjoin_builder::
 * consume_warp (hashjoin.cu:41-60): next[s] = heads[b]; heads[b] = s. */
int orc_ct_build(orc_chained_table *t, const int32_t *keys, const int32_t *payload,
                 uint64_t n, uint32_t log_buckets) {
    uint64_t nb = (uint64_t)1 << log_buckets;
    t->heads = (int32_t *)malloc(nb * sizeof(int32_t));
    t->next = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
    if (!t->heads || !t->next) { free(t->heads); free(t->next); return -1; }
    memset(t->heads, 0xff, nb * sizeof(int32_t));
    t->keys = keys;
    t->payload = payload;
    t->log_buckets = log_buckets;
    t->n = n;
    uint32_t mask = (uint32_t)(nb - 1);
    for (uint64_t s = 0; s < n; ++s) {
        uint32_t b = orc_fmix32((uint32_t)keys[s]) & mask;
        t->next[s] = t->heads[b];
        t->heads[b] = (int32_t)s;
    }
    return 0;
}

void orc_ct_free(orc_chained_table *t) {
    free(t->heads);
    free(t->next);
    t->heads = t->next = NULL;
}

/* Probe of sum_hashjoin_select_cpu (main8.cu:93-100) / hashjoin::consume_warp
 * (hashjoin.cu:135-182): walk the chain of fmix32(key) & mask and count key
 * equality; extended with the payload sum and the pair digest. */
void orc_ct_probe(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                  const int32_t *rk, const int32_t *rv, uint64_t nr, uint64_t r_base,
                  orc_join_result *res) {
    uint32_t mask = (uint32_t)(((uint64_t)1 << t->log_buckets) - 1);
    uint64_t cnt = 0, dsum = 0, dxor = 0;
    int64_t sum = 0;
    for (uint64_t r = 0; r < nr; ++r) {
        if (pred && !(pred[r] <= bound)) continue;
        int32_t k = rk[r];
        int32_t cur = t->heads[orc_fmix32((uint32_t)k) & mask];
        while (cur >= 0) {
            if (t->keys[cur] == k) {
                ++cnt;
                sum += (int64_t)(rv ? rv[r] : 0) + (int64_t)(t->payload ? t->payload[cur] : 0);
                uint64_t d = orc_mix64(((r_base + r) << 32) | (uint64_t)(uint32_t)cur);
                dsum += d;
                dxor ^= d;
            }
            cur = t->next[cur];
        }
    }
    res->count = cnt;
    res->sum = sum;
    res->digest_sum = dsum;
    res->digest_xor = dxor;
    res->count32 = (int32_t)(uint32_t)cnt;
}

typedef struct {
    const orc_chained_table *t;
    const int32_t *pred, *rk, *rv;
    int32_t bound;
    uint64_t lo, hi;
    orc_join_result res;
} probe_task;

static void *probe_worker(void *p) {
    probe_task *k = (probe_task *)p;
    orc_ct_probe(k->t, k->pred ? k->pred + k->lo : NULL, k->bound, k->rk + k->lo,
                 k->rv ? k->rv + k->lo : NULL, k->hi - k->lo, k->lo, &k->res);
    return NULL;
}

void orc_ct_probe_mt(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                     const int32_t *rk, const int32_t *rv, uint64_t nr,
                     orc_join_result *res, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    probe_task tk[256];
    for (int i = 0; i < nthreads; ++i) {
        tk[i].t = t; tk[i].pred = pred; tk[i].rk = rk; tk[i].rv = rv; tk[i].bound = bound;
        tk[i].lo = nr * (uint64_t)i / (uint64_t)nthreads;
        tk[i].hi = nr * (uint64_t)(i + 1) / (uint64_t)nthreads;
        pthread_create(&th[i], NULL, probe_worker, &tk[i]);
    }
    memset(res, 0, sizeof(*res));
    for (int i = 0; i < nthreads; ++i) {
        pthread_join(th[i], NULL);
        res->count += tk[i].res.count;
        res->sum += tk[i].res.sum;
        res->digest_sum += tk[i].res.digest_sum;
        res->digest_xor ^= tk[i].res.digest_xor;
    }
    res->count32 = (int32_t)(uint32_t)res->count;
}

/* Multi-threaded build of the same chained table: thread i inserts rows
 * [n*i/T, n*(i+1)/T) with an atomic exchange on the bucket head, which is
 * the reference's own device build (hashjoin.cu:54-55: next[s] =
 * atomicExch(&first[b], s)).  The chains hold the same rows as the serial
 * build, in some order; probe results do not depend on that order. */
typedef struct {
    orc_chained_table *t;
    uint64_t lo, hi;
} build_task;

static void *build_worker(void *p) {
    build_task *k = (build_task *)p;
    orc_chained_table *t = k->t;
    const uint32_t mask = (uint32_t)(((uint64_t)1 << t->log_buckets) - 1);
    for (uint64_t s = k->lo; s < k->hi; ++s) {
        const uint32_t b = orc_fmix32((uint32_t)t->keys[s]) & mask;
        t->next[s] = __atomic_exchange_n(&t->heads[b], (int32_t)s, __ATOMIC_RELAXED);
    }
    return NULL;
}

int orc_ct_build_mt(orc_chained_table *t, const int32_t *keys, const int32_t *payload,
                    uint64_t n, uint32_t log_buckets, int nthreads) {
    if (orc_ct_build(t, keys, payload, 0, log_buckets)) return -1;   /* heads only */
    free(t->next);
    t->next = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
    if (!t->next) { orc_ct_free(t); return -1; }
    t->n = n;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    build_task tk[256];
    for (int i = 0; i < nthreads; ++i) {
        tk[i].t = t;
        tk[i].lo = n * (uint64_t)i / (uint64_t)nthreads;
        tk[i].hi = n * (uint64_t)(i + 1) / (uint64_t)nthreads;
        pthread_create(&th[i], NULL, build_worker, &tk[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    return 0;
}

/* Join against a build side whose keys are distinct and lie in [0, kmax),
 * by direct addressing (no hash table): for every probe row r (optionally
 * pred[r] <= bound) whose key is a build key s: count += 1, sum += rv[r] +
 * payload[s].  For the C2 inputs (build keys a permutation of [0, 2^24))
 * this is the closed form COUNT = #{r : R.k < 2^24}, SUM = sum of R.v +
 * w[perm^-1(R.k)] over those rows.  Returns -1 (and computes nothing) when
 * a build key is out of range or repeated. */
typedef struct {
    const int32_t *pred, *rk, *rv;
    int32_t bound;
    const int32_t *w;      /* payload by key */
    const uint8_t *has;    /* key present */
    uint32_t kmax;
    uint64_t lo, hi;
    uint64_t count;
    int64_t sum;
} dense_task;

static void *dense_worker(void *p) {
    dense_task *k = (dense_task *)p;
    uint64_t c = 0;
    int64_t s = 0;
    for (uint64_t r = k->lo; r < k->hi; ++r) {
        if (k->pred && !(k->pred[r] <= k->bound)) continue;
        const uint32_t key = (uint32_t)k->rk[r];
        if (key < k->kmax && k->has[key]) {
            ++c;
            s += (int64_t)(k->rv ? k->rv[r] : 0) + (int64_t)k->w[key];
        }
    }
    k->count = c;
    k->sum = s;
    return NULL;
}

int orc_join_dense(const int32_t *sk, const int32_t *sw, uint64_t ns, uint32_t kmax, const int32_t *pred,
                   int32_t bound, const int32_t *rk, const int32_t *rv, uint64_t nr, uint64_t *count,
                   int64_t *sum, int nthreads) {
    int32_t *w = (int32_t *)calloc(kmax ? kmax : 1, sizeof(int32_t));
    uint8_t *has = (uint8_t *)calloc(kmax ? kmax : 1, 1);
    if (!w || !has) { free(w); free(has); return -1; }
    for (uint64_t i = 0; i < ns; ++i) {
        const uint32_t key = (uint32_t)sk[i];
        if (key >= kmax || has[key]) { free(w); free(has); return -1; }
        has[key] = 1;
        w[key] = sw ? sw[i] : 0;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    dense_task tk[256];
    for (int i = 0; i < nthreads; ++i) {
        tk[i].pred = pred; tk[i].rk = rk; tk[i].rv = rv; tk[i].bound = bound;
        tk[i].w = w; tk[i].has = has; tk[i].kmax = kmax;
        tk[i].lo = nr * (uint64_t)i / (uint64_t)nthreads;
        tk[i].hi = nr * (uint64_t)(i + 1) / (uint64_t)nthreads;
        pthread_create(&th[i], NULL, dense_worker, &tk[i]);
    }
    uint64_t c = 0;
    int64_t s = 0;
    for (int i = 0; i < nthreads; ++i) {
        pthread_join(th[i], NULL);
        c += tk[i].count;
        s += tk[i].sum;
    }
    free(w);
    free(has);
    *count = c;
    *sum = s;
    return 0;
}

typedef struct {
    const orc_chained_table *t;
    const int32_t *pred, *rk, *rv;
    int32_t bound;
    uint64_t lo, hi;
    uint64_t *cnt;
    int64_t *sum;
    uint32_t ngroups;
} pgb_task;

static void *pgb_worker(void *p) {
    pgb_task *k = (pgb_task *)p;
    orc_ct_probe_groupby(k->t, k->pred ? k->pred + k->lo : NULL, k->bound, k->rk + k->lo, k->rv + k->lo,
                         k->hi - k->lo, k->cnt, k->sum, k->ngroups);
    return NULL;
}

/* orc_ct_probe_groupby on nthreads threads: per-thread group arrays over
 * row ranges, added up at the end. */
int orc_ct_probe_groupby_mt(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                            const int32_t *rk, const int32_t *rv, uint64_t nr,
                            uint64_t *cnt, int64_t *sum, uint32_t ngroups, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    uint64_t *pc = (uint64_t *)malloc((size_t)nthreads * ngroups * sizeof(uint64_t));
    int64_t *ps = (int64_t *)malloc((size_t)nthreads * ngroups * sizeof(int64_t));
    if (!pc || !ps) { free(pc); free(ps); return -1; }
    pthread_t th[256];
    pgb_task tk[256];
    for (int i = 0; i < nthreads; ++i) {
        tk[i].t = t; tk[i].pred = pred; tk[i].rk = rk; tk[i].rv = rv; tk[i].bound = bound;
        tk[i].lo = nr * (uint64_t)i / (uint64_t)nthreads;
        tk[i].hi = nr * (uint64_t)(i + 1) / (uint64_t)nthreads;
        tk[i].cnt = pc + (size_t)i * ngroups;
        tk[i].sum = ps + (size_t)i * ngroups;
        tk[i].ngroups = ngroups;
        pthread_create(&th[i], NULL, pgb_worker, &tk[i]);
    }
    memset(cnt, 0, ngroups * sizeof(uint64_t));
    memset(sum, 0, ngroups * sizeof(int64_t));
    for (int i = 0; i < nthreads; ++i) {
        pthread_join(th[i], NULL);
        for (uint32_t g = 0; g < ngroups; ++g) {
            cnt[g] += tk[i].cnt[g];
            sum[g] += tk[i].sum[g];
        }
    }
    free(pc);
    free(ps);
    return 0;
}

void orc_ct_probe_groupby(const orc_chained_table *t, const int32_t *pred, int32_t bound,
                          const int32_t *rk, const int32_t *rv, uint64_t nr,
                          uint64_t *cnt, int64_t *sum, uint32_t ngroups) {
    memset(cnt, 0, ngroups * sizeof(uint64_t));
    memset(sum, 0, ngroups * sizeof(int64_t));
    uint32_t mask = (uint32_t)(((uint64_t)1 << t->log_buckets) - 1);
    for (uint64_t r = 0; r < nr; ++r) {
        if (pred && !(pred[r] <= bound)) continue;
        int32_t k = rk[r];
        int32_t cur = t->heads[orc_fmix32((uint32_t)k) & mask];
        while (cur >= 0) {
            if (t->keys[cur] == k) {
                uint32_t g = (uint32_t)t->payload[cur];
                if (g < ngroups) {
                    cnt[g] += 1;
                    sum[g] += rv[r];
                }
            }
            cur = t->next[cur];
        }
    }
}

void orc_groupby(const int32_t *key, const int32_t *val, uint64_t n,
                 uint64_t *cnt, int64_t *sum, uint32_t ngroups) {
    memset(cnt, 0, ngroups * sizeof(uint64_t));
    memset(sum, 0, ngroups * sizeof(int64_t));
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t g = (uint32_t)key[i];
        if (g < ngroups) {
            cnt[g] += 1;
            sum[g] += val[i];
        }
    }
}

/* sum_selfjoin_select_cpu, main9.cu:63-69 (hist over the filtered values, then
 * sum of squares), generalised to values in [0, vmax]. */
int64_t orc_selfjoin_count(const int32_t *x, uint64_t n, int32_t bound, uint32_t vmax) {
    uint64_t *hist = (uint64_t *)calloc((size_t)vmax + 1, sizeof(uint64_t));
    if (!hist) return -1;
    for (uint64_t j = 0; j < n; ++j)
        if (x[j] <= bound && x[j] >= 0 && (uint32_t)x[j] <= vmax) ++hist[x[j]];
    int64_t s = 0;
    for (uint32_t v = 0; v <= vmax; ++v) s += (int64_t)(hist[v] * hist[v]);
    free(hist);
    return s;
}

/* split.cu:54-118: the destination of a row is hash3 of its value (49-51);
 * consume_warp only ballots destinations 0 .. num_of_parents-1, so a row
 * whose destination has no parent is never written. */
void orc_split(const int32_t *pred, int32_t bound, const int32_t *const *cols, uint32_t ncols, uint64_t n,
               uint32_t nparents, int32_t *const *out, uint64_t *counts) {
    for (uint32_t p = 0; p < nparents; ++p) counts[p] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (pred && pred[i] > bound) continue;
        const uint32_t d = (uint32_t)(cols[0][i] & 1);
        if (d >= nparents) continue;
        for (uint32_t c = 0; c < ncols; ++c) out[d * ncols + c][counts[d]] = cols[c][i];
        ++counts[d];
    }
}
