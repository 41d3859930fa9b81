/*
 * gen.c — seeded synthetic tables for the GROUP BY workloads (host, C).
 *
 * Every value is a pure function of (seed, column stream, row index): a
 * counter-based SplitMix64 draw, so any slice of any table can be produced
 * independently, by any number of threads, and again later with the same
 * bits.  Distributions:
 *   plain   : k BIGINT U[0,1024), v BIGINT U[-1e6, 1e6]
 *   filtered: region INT Zipf(s=1.1, n=10000) over rank-permuted ids,
 *             ts DATETIME (us) uniform over 30 days from 2026-01-01 UTC,
 *             amount DECIMAL(18,2) cents = min(round(exp(8 + 2 z)), 1e12),
 *             price DOUBLE U[0,1000)
 *   highcard: user_id BIGINT U[0,2e7), bucket TINYINT U[0,16),
 *             v INT U[0,1e5), w DOUBLE N(0,1)
 *   batched : per table, rows = round(exp(U(ln 1e4, ln 2e7))),
 *             card = round(exp(U(ln 4, ln 1e6))), key INT U[0,card),
 *             value BIGINT U[-1e9, 1e9], filter coin flip
 * Normal draws use Box-Muller on two uniforms of the same row.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GEN_MAX_THREADS 64

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t row) {
  return mix64(mix64(seed ^ (stream * 0x9e3779b97f4a7c15ull)) + row * 0xd1b54a32d192ed03ull);
}
static inline double unif(uint64_t seed, uint64_t stream, uint64_t row) {
  return (double)(draw(seed, stream, row) >> 11) * 0x1.0p-53;
}
static inline double normal(uint64_t seed, uint64_t stream, uint64_t row) {
  double u1 = unif(seed, stream, row), u2 = unif(seed, stream + 1000, row);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

/* ---- Zipf(s, n) inverse CDF over rank-permuted ids ------------------------- */
typedef struct {
  uint64_t seed;
  uint32_t n;
  double s;
  double* cdf;
  int32_t* ids;
} zipf_t;

static pthread_mutex_t g_zipf_lock = PTHREAD_MUTEX_INITIALIZER;
static zipf_t g_zipf;

static const zipf_t* zipf_get(uint64_t seed, uint32_t n, double s) {
  pthread_mutex_lock(&g_zipf_lock);
  if (!g_zipf.cdf || g_zipf.seed != seed || g_zipf.n != n || g_zipf.s != s) {
    free(g_zipf.cdf);
    free(g_zipf.ids);
    g_zipf.cdf = malloc(n * sizeof(double));
    g_zipf.ids = malloc(n * sizeof(int32_t));
    double h = 0;
    for (uint32_t k = 0; k < n; k++) {
      h += pow((double)(k + 1), -s);
      g_zipf.cdf[k] = h;
    }
    for (uint32_t k = 0; k < n; k++) g_zipf.cdf[k] /= h;
    g_zipf.cdf[n - 1] = 1.0;
    for (uint32_t k = 0; k < n; k++) g_zipf.ids[k] = (int32_t)k;
    for (uint32_t k = n - 1; k > 0; k--) { /* Fisher-Yates, fixed seed */
      uint32_t j = (uint32_t)(draw(seed, 77, k) % (k + 1));
      int32_t t = g_zipf.ids[k];
      g_zipf.ids[k] = g_zipf.ids[j];
      g_zipf.ids[j] = t;
    }
    g_zipf.seed = seed;
    g_zipf.n = n;
    g_zipf.s = s;
  }
  pthread_mutex_unlock(&g_zipf_lock);
  return &g_zipf;
}

static inline int32_t zipf_draw(const zipf_t* z, double u) {
  uint32_t lo = 0, hi = z->n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (z->cdf[mid] > u) hi = mid; else lo = mid + 1;
  }
  return z->ids[lo];
}

/* ---- threaded fill ---------------------------------------------------------- */
typedef struct {
  int kind;
  uint64_t seed, row0, n;
  void* out[4];
  const zipf_t* z;
  int64_t p0, p1; /* batched: card */
} job_t;

#define T0_US 1767225600000000ll /* 2026-01-01 00:00:00 UTC */
#define DAY_US 86400000000ll

static void fill(const job_t* j, uint64_t lo, uint64_t hi) {
  for (uint64_t i = lo; i < hi; i++) {
    uint64_t r = j->row0 + i;
    switch (j->kind) {
      case 1:
        ((int64_t*)j->out[0])[i] = (int64_t)(unif(j->seed, 1, r) * 1024.0);
        ((int64_t*)j->out[1])[i] = (int64_t)(unif(j->seed, 2, r) * 2000001.0) - 1000000;
        break;
      case 2: {
        ((int32_t*)j->out[0])[i] = zipf_draw(j->z, unif(j->seed, 1, r));
        ((int64_t*)j->out[1])[i] = T0_US + (int64_t)(unif(j->seed, 2, r) * (double)(30 * DAY_US));
        double a = exp(8.0 + 2.0 * normal(j->seed, 3, r));
        int64_t cents = a >= 1e12 ? 1000000000000ll : llround(a);
        ((int64_t*)j->out[2])[i] = cents;
        ((double*)j->out[3])[i] = unif(j->seed, 4, r) * 1000.0;
        break;
      }
      case 3:
        ((int64_t*)j->out[0])[i] = (int64_t)(unif(j->seed, 1, r) * 2e7);
        ((int8_t*)j->out[1])[i] = (int8_t)(unif(j->seed, 2, r) * 16.0);
        ((int32_t*)j->out[2])[i] = (int32_t)(unif(j->seed, 3, r) * 1e5);
        ((double*)j->out[3])[i] = normal(j->seed, 4, r);
        break;
      case 4:
        ((int32_t*)j->out[0])[i] = (int32_t)(unif(j->seed, 1, r) * (double)j->p0);
        ((int64_t*)j->out[1])[i] = (int64_t)(unif(j->seed, 2, r) * 2000000001.0) - 1000000000;
        break;
      default: break;
    }
  }
}

typedef struct {
  const job_t* j;
  uint64_t lo, hi;
} span_t;

static void* fill_main(void* a) {
  span_t* s = a;
  fill(s->j, s->lo, s->hi);
  return NULL;
}

static void run(const job_t* j, int threads) {
  if (threads < 1) threads = 1;
  if (threads > GEN_MAX_THREADS) threads = GEN_MAX_THREADS;
  if (j->n < 65536 || threads == 1) {
    fill(j, 0, j->n);
    return;
  }
  pthread_t th[GEN_MAX_THREADS];
  span_t sp[GEN_MAX_THREADS];
  for (int t = 0; t < threads; t++) {
    sp[t].j = j;
    sp[t].lo = j->n * t / threads;
    sp[t].hi = j->n * (t + 1) / threads;
    pthread_create(&th[t], NULL, fill_main, &sp[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
}

/* Rows [row0, row0+n) of the plain table. */
void tdbg_gen_plain(uint64_t seed, uint64_t row0, uint64_t n, int64_t* k, int64_t* v, int threads) {
  job_t j = {1, seed, row0, n, {k, v, 0, 0}, 0, 0, 0};
  run(&j, threads);
}

void tdbg_gen_filtered(uint64_t seed, uint64_t row0, uint64_t n, int32_t* region, int64_t* ts,
                       int64_t* amount, double* price, int threads) {
  job_t j = {2, seed, row0, n, {region, ts, amount, price}, zipf_get(seed, 10000, 1.1), 0, 0};
  run(&j, threads);
}

void tdbg_gen_highcard(uint64_t seed, uint64_t row0, uint64_t n, int64_t* user_id, int8_t* bucket,
                       int32_t* v, double* w, int threads) {
  job_t j = {3, seed, row0, n, {user_id, bucket, v, w}, 0, 0, 0};
  run(&j, threads);
}

/* Shape of batched table `t`: row count, key cardinality, filter flag. */
void tdbg_gen_batched_shape(uint64_t seed, uint32_t t, uint64_t* rows, int64_t* card, int32_t* filtered) {
  double lr = log(1e4) + unif(seed, 101, t) * (log(2e7) - log(1e4));
  double lc = log(4.0) + unif(seed, 102, t) * (log(1e6) - log(4.0));
  *rows = (uint64_t)llround(exp(lr));
  *card = (int64_t)llround(exp(lc));
  *filtered = unif(seed, 103, t) < 0.5;
}

void tdbg_gen_batched(uint64_t seed, uint32_t t, uint64_t row0, uint64_t n, int32_t* key, int64_t* value,
                      int threads) {
  uint64_t rows;
  int64_t card;
  int32_t f;
  tdbg_gen_batched_shape(seed, t, &rows, &card, &f);
  job_t j = {4, mix64(seed + 0x1000 + t), row0, n, {key, value, 0, 0}, 0, card, 0};
  run(&j, threads);
}

/* A batched-workload-shaped table with a given key cardinality: key INT
 * U[0, card), value BIGINT U[-1e9, 1e9]. */
void tdbg_gen_keyed(uint64_t seed, int64_t card, uint64_t row0, uint64_t n, int32_t* key, int64_t* value,
                    int threads) {
  job_t j = {4, mix64(seed + 0x2000), row0, n, {key, value, 0, 0}, 0, card, 0};
  run(&j, threads);
}

/* The Zipf probability mass of rank k (1-based), for distribution tests. */
double tdbg_gen_zipf_pmf(uint32_t n, double s, uint32_t k) {
  double h = 0;
  for (uint32_t i = 1; i <= n; i++) h += pow((double)i, -s);
  return pow((double)k, -s) / h;
}
