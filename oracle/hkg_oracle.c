/*
 * TEST INFRASTRUCTURE ONLY -- the oracle, i.e. the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and never as the thing measured or shipped.
 *
 * Plain-C restatement of the reference's FP32 arithmetic on the exact k-NN
 * path (built with -ffp-contract=off so no product is fused unless a
 * function says so), plus the workload generator specification.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   orc_ed2, orc_paa, orc_sax_word, orc_lb_sax, orc_seg_mean_std,
 *   orc_lb_eapca(use_fma=1)   bit-equal to the reference functions compiled
 *                             from their own sources (oracle/_ref), on random
 *                             and edge inputs; plus closed forms/invariants
 *   orc_breakpoints256        float32(Phi^-1(i/256)) (closed form) and the
 *                             reference's own table
 *   orc_knn_bruteforce        the reference CLI's answers (serial, SIMS and
 *                             PSCAN paths) and brute force in FP64
 *   orc_knn_mode              the reference CLI's answers with --epsilon /
 *                             --approx-stop-condition / --nprobes on the
 *                             committed golden indexes
 *   generator                 a specification, not a reference function: its
 *                             invariants (z-normalisation, determinism) and
 *                             bit-equality with the product's HIP generator
 */
#define _GNU_SOURCE /* qsort_r */
#include "hkg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* generator specification                                                   */
/* ------------------------------------------------------------------------ */

uint64_t orc_mix64(uint64_t x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

/* one 64-bit draw: counter (stream, j) with j < 2^20; mix64 is a bijection,
   so distinct counters under one seed never collide */
uint64_t orc_draw(uint64_t seed, uint64_t stream, uint64_t j)
{
  uint64_t key = orc_mix64(seed ^ 0x9E3779B97F4A7C15ULL);
  uint64_t c = (stream << 20) | (j & 0xFFFFFULL);
  return orc_mix64(key ^ orc_mix64(c));
}

double orc_uniform(uint64_t bits)
{
  return ((double)(bits >> 11) + 0.5) * 0x1.0p-53; /* in (0, 1) */
}

/* natural log from basic IEEE operations only (so CPU and GPU agree):
   u = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) */
double orc_log(double u)
{
  uint64_t b;
  memcpy(&b, &u, 8);
  int e = (int)((b >> 52) & 0x7FF) - 1022;
  uint64_t mb = (b & 0x000FFFFFFFFFFFFFULL) | (1022ULL << 52);
  double m;
  memcpy(&m, &mb, 8);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double t = 1.0 / 23.0;
  t = t * s2 + 1.0 / 21.0;
  t = t * s2 + 1.0 / 19.0;
  t = t * s2 + 1.0 / 17.0;
  t = t * s2 + 1.0 / 15.0;
  t = t * s2 + 1.0 / 13.0;
  t = t * s2 + 1.0 / 11.0;
  t = t * s2 + 1.0 / 9.0;
  t = t * s2 + 1.0 / 7.0;
  t = t * s2 + 1.0 / 5.0;
  t = t * s2 + 1.0 / 3.0;
  t = t * s2 + 1.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  return (double)e * ln2_hi + ((2.0 * s) * t + (double)e * ln2_lo);
}

double orc_crsqrt(double x) { return sqrt(x); } /* IEEE correctly rounded */

/* sin/cos of a in [0, pi/2) by nested Taylor factors */
static void orc_sincos_q(double a, double *sn, double *cs)
{
  double a2 = a * a;
  double s = 1.0, c = 1.0;
  for (int k = 13; k >= 1; --k) {
    s = 1.0 - (a2 / (double)((2 * k) * (2 * k + 1))) * s;
    c = 1.0 - (a2 / (double)((2 * k - 1) * (2 * k))) * c;
  }
  *sn = a * s;
  *cs = c;
}

void orc_normal_pair(uint64_t seed, uint64_t stream, uint64_t pair, double *z0, double *z1)
{
  double u1 = orc_uniform(orc_draw(seed, stream, 2 * pair));
  double u2 = orc_uniform(orc_draw(seed, stream, 2 * pair + 1));
  double r = orc_crsqrt(-2.0 * orc_log(u1));
  double t4 = u2 * 4.0;
  int q = (int)t4;
  double a = (t4 - (double)q) * 1.57079632679489661923;
  double sn, cs;
  orc_sincos_q(a, &sn, &cs);
  double c, s;
  switch (q & 3) {
  case 0: c = cs; s = sn; break;
  case 1: c = -sn; s = cs; break;
  case 2: c = -cs; s = -sn; break;
  default: c = sn; s = -cs; break;
  }
  *z0 = r * c;
  *z1 = r * s;
}

/* z-normalise v[0..L) (population sigma) into float out */
static void orc_znorm_to_float(const double *v, uint32_t L, float *out)
{
  double sum = 0.0;
  for (uint32_t t = 0; t < L; ++t) sum = sum + v[t];
  double mu = sum / (double)L;
  double ss = 0.0;
  for (uint32_t t = 0; t < L; ++t) {
    double d = v[t] - mu;
    ss = ss + d * d;
  }
  double sd = orc_crsqrt(ss / (double)L);
  for (uint32_t t = 0; t < L; ++t) out[t] = (float)((v[t] - mu) / sd);
}

static void orc_one_series(int kind, uint64_t seed, uint64_t id, uint32_t L, double *v, float *out)
{
  for (uint32_t p = 0; p < L / 2; ++p) orc_normal_pair(seed, id, p, &v[2 * p], &v[2 * p + 1]);
  if (kind == 0) {
    double s = 0.0;
    for (uint32_t t = 0; t < L; ++t) {
      s = s + v[t];
      v[t] = s;
    }
  } else {
    double n2 = 0.0;
    for (uint32_t t = 0; t < L; ++t) n2 = n2 + v[t] * v[t];
    double nrm = orc_crsqrt(n2);
    for (uint32_t t = 0; t < L; ++t) v[t] = v[t] / nrm;
  }
  orc_znorm_to_float(v, L, out);
}

struct gen_job {
  int kind;
  uint64_t seed, first, n;
  uint32_t L;
  float *out;
};

static void *orc_gen_worker(void *p)
{
  struct gen_job *j = (struct gen_job *)p;
  double *v = malloc(sizeof(double) * j->L);
  for (uint64_t i = 0; i < j->n; ++i)
    orc_one_series(j->kind, j->seed, j->first + i, j->L, v, j->out + i * j->L);
  free(v);
  return NULL;
}

void orc_gen_series(int kind, uint64_t seed, uint64_t first_id, uint64_t n, uint32_t L,
                    float *out, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > n) nthreads = n ? (int)n : 1;
  pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
  struct gen_job *jobs = malloc(sizeof(struct gen_job) * nthreads);
  uint64_t per = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    uint64_t b = (uint64_t)t * per, e = b + per;
    if (b > n) b = n;
    if (e > n) e = n;
    jobs[t] = (struct gen_job){kind, seed, first_id + b, e - b, L, out + b * L};
    pthread_create(&th[t], NULL, orc_gen_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
}

void orc_gen_noisy_queries(int kind, uint64_t data_seed, uint64_t N, uint64_t query_seed,
                           uint32_t nq, uint32_t L, const float *sigmas,
                           float *out, uint64_t *src_ids)
{
  double *v = malloc(sizeof(double) * L);
  float *x = malloc(sizeof(float) * L);
  for (uint32_t i = 0; i < nq; ++i) {
    uint64_t src = orc_draw(query_seed, i, 0xFFFFF) % N;
    if (src_ids) src_ids[i] = src;
    orc_one_series(kind, data_seed, src, L, v, x);
    double sg = (double)sigmas[i];
    for (uint32_t p = 0; p < L / 2; ++p) {
      double z0, z1;
      orc_normal_pair(query_seed, i, p, &z0, &z1);
      v[2 * p] = (double)x[2 * p] + sg * z0;
      v[2 * p + 1] = (double)x[2 * p + 1] + sg * z1;
    }
    orc_znorm_to_float(v, L, out + (uint64_t)i * L);
  }
  free(v);
  free(x);
}

/* ------------------------------------------------------------------------ */
/* reference arithmetic                                                      */
/* ------------------------------------------------------------------------ */

/* ts.c:83-109: per 8-chunk, d_j = (q-x)^2 rounded separately, chunk sum as
   ((d0+d1)+(d2+d3)) + ((d4+d5)+(d6+d7)) (the two hadds), accumulated
   sequentially; the loop runs while acc < bound, checked before each chunk */
float orc_ed2(const float *q, const float *x, int L, float bound)
{
  float acc = 0.0f;
  for (int i = 0; i < L && acc < bound; i += 8) {
    float d[8];
    for (int j = 0; j < 8; ++j) {
      float t = q[i + j] - x[i + j];
      d[j] = t * t;
    }
    float s = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7]));
    acc = acc + s;
  }
  return acc;
}

double orc_ed2_f64(const float *q, const float *x, int L)
{
  double acc = 0.0;
  for (int i = 0; i < L; ++i) {
    double t = (double)q[i] - (double)x[i];
    acc += t * t;
  }
  return acc;
}

/* sax.c:46-56: sequential sum per segment, then divide by the segment size */
void orc_paa(const float *x, int L, int w, float *out)
{
  int per = L / w;
  for (int s = 0; s < w; ++s) {
    float a = 0.0f;
    for (int i = 0; i < per; ++i) a = a + x[s * per + i];
    out[s] = a / (float)per;
  }
}

/* Acklam-style rational start refined by Newton steps on the FP64 normal
   CDF (erfc), giving Phi^-1 to ~1e-16; the card-256 iSAX breakpoints
   (include/sax_breakpoints.h, slice at offset 255*254/2) are
   float32(Phi^-1(i/256)), i = 1..255 */
static double orc_norm_cdf(double x) { return 0.5 * erfc(-x / 1.41421356237309504880); }

void orc_breakpoints256(float *out255)
{
  for (int i = 1; i <= 255; ++i) {
    double p = (double)i / 256.0;
    if (i == 128) { /* Phi^-1(1/2) = 0 exactly */
      out255[i - 1] = 0.0f;
      continue;
    }
    /* start: logit-based approximation, then Newton/Halley on Phi(x) - p */
    double x = (p < 0.5 ? -1.0 : 1.0) * sqrt(-2.0 * log(p < 0.5 ? p : 1.0 - p));
    x = x - ((p < 0.5 ? -1.0 : 1.0) * (2.515517 + 0.802853 * fabs(x) + 0.010328 * x * x) /
             (1.0 + 1.432788 * fabs(x) + 0.189269 * x * x + 0.001308 * fabs(x) * x * x));
    for (int it = 0; it < 60; ++it) {
      double f = orc_norm_cdf(x) - p;
      double dens = exp(-0.5 * x * x) / 2.50662827463100050242;
      double step = f / dens;
      double nx = x - step / (1.0 + 0.5 * x * step);
      if (nx == x) break;
      x = nx;
    }
    out255[i - 1] = (float)x;
  }
}

/* sax.c:86-135 with the bsearch comparator at sax.c:28-41: the symbol is the
   number of breakpoints strictly below the PAA value */
void orc_sax_word(const float *paa, int w, const float *bp, uint8_t *out)
{
  for (int s = 0; s < w; ++s) {
    int lo = 0, hi = 255; /* first index with bp >= p */
    float p = paa[s];
    while (lo < hi) {
      int mid = (lo + hi) / 2;
      if (bp[mid] < p) lo = mid + 1;
      else hi = mid;
    }
    out[s] = (uint8_t)lo;
  }
}

/* sax.c:310-479 (16 segments, full 8-bit cardinality words) */
float orc_lb_sax(const float *paa, const uint8_t *sax, int L, const float *bp)
{
  float t[16];
  for (int i = 0; i < 16; ++i) {
    int v = sax[i];
    float lo = v == 0 ? -2000000.0f : bp[v - 1];
    float hi = v == 255 ? 2000000.0f : bp[v];
    float p = paa[i];
    if (lo > p) t[i] = (lo - p) * (lo - p);
    else if (hi < p) t[i] = (hi - p) * (hi - p);
    else t[i] = 0.0f;
  }
  float w[8];
  for (int i = 0; i < 8; ++i) w[i] = t[i] + t[i + 8];
  float s = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
  float ratio = (float)(L / 16); /* hercules_index.c:106-107, integer division */
  return s * ratio;
}

static float orc_tree8(const float *v)
{
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

/* calc_utils.c:48-112 over [s, e): 8-chunks from s, the last zero-padded */
void orc_seg_mean_std(const float *x, int s, int e, float *mu, float *sd)
{
  int n = e - s;
  float sx = 0.0f, sq = 0.0f;
  const float *p = x + s;
  int d = n;
  while (d > 0) {
    float c[8], c2[8];
    for (int j = 0; j < 8; ++j) c[j] = j < d ? p[j] : 0.0f;
    for (int j = 0; j < 8; ++j) c2[j] = c[j] * c[j];
    sx = sx + orc_tree8(c);
    sq = sq + orc_tree8(c2);
    p += 8;
    d -= 8;
  }
  *mu = sx / (float)n;
  sq = sq - (sx * sx) / (float)n;
  *sd = (float)sqrt((double)(sq / (float)n));
}

/* hercules_node.c:879-924; ind = {max mean, min mean, max sd, min sd} per
   segment.  use_fma=1 reproduces the three contractions GCC makes under an
   FMA-capable -march (temp_dist += temp*temp twice, sum += temp_dist*len). */
float orc_lb_eapca(const float *q, int nseg, const int16_t *seg_end, const float *ind, int use_fma)
{
  float sum = 0.0f;
  int start = 0;
  for (int i = 0; i < nseg; ++i) {
    int end = seg_end[i];
    float mu, sd;
    orc_seg_mean_std(q, start, end, &mu, &sd);
    const float *g = ind + 4 * i;
    float td = 0.0f;
    if ((sd - g[2]) * (sd - g[3]) > 0.0f) {
      float t = (float)fmin(fabs((double)(sd - g[2])), fabs((double)(sd - g[3])));
      td = use_fma ? fmaf(t, t, td) : td + t * t;
    }
    if ((mu - g[0]) * (mu - g[1]) > 0.0f) {
      float t = (float)fmin(fabs((double)(mu - g[0])), fabs((double)(mu - g[1])));
      td = use_fma ? fmaf(t, t, td) : td + t * t;
    }
    float len = (float)(end - start);
    sum = use_fma ? fmaf(td, len, sum) : sum + td * len;
    start = end;
  }
  return sum;
}

/* ------------------------------------------------------------------------ */
/* brute-force answers                                                       */
/* ------------------------------------------------------------------------ */

/* hercules_query_engine.c:4034-4082 semantics (distinct distances, strict <
   at the call sites) with ids visited in ascending order, so the first id
   seen for a distance is the smallest */
static void orc_topk_insert(float *d, uint64_t *id, uint32_t *cur, uint32_t k, float nd, uint64_t nid)
{
  for (uint32_t i = 0; i < *cur; ++i)
    if (d[i] == nd) return;
  uint32_t pos;
  if (*cur == k) pos = k - 1;
  else pos = (*cur)++;
  d[pos] = nd;
  id[pos] = nid;
  while (pos > 0 && d[pos - 1] > d[pos]) {
    float td = d[pos]; d[pos] = d[pos - 1]; d[pos - 1] = td;
    uint64_t ti = id[pos]; id[pos] = id[pos - 1]; id[pos - 1] = ti;
    --pos;
  }
}

struct bf_job {
  const float *X, *Q;
  uint64_t b, e;
  int L;
  uint32_t nq, k;
  float *d;      /* nq*k */
  uint64_t *id;  /* nq*k */
  uint32_t *cur; /* nq */
};

static void *orc_bf_worker(void *p)
{
  struct bf_job *j = (struct bf_job *)p;
  const uint64_t BLK = 2048;
  for (uint64_t b0 = j->b; b0 < j->e; b0 += BLK) {
    uint64_t b1 = b0 + BLK < j->e ? b0 + BLK : j->e;
    for (uint32_t qi = 0; qi < j->nq; ++qi) {
      const float *q = j->Q + (uint64_t)qi * j->L;
      float *dd = j->d + (uint64_t)qi * j->k;
      uint64_t *ii = j->id + (uint64_t)qi * j->k;
      uint32_t *cur = &j->cur[qi];
      for (uint64_t r = b0; r < b1; ++r) {
        float kth = *cur == j->k ? dd[j->k - 1] : FLT_MAX;
        float dist = orc_ed2(q, j->X + r * j->L, j->L, kth);
        if (dist < kth) orc_topk_insert(dd, ii, cur, j->k, dist, r);
      }
    }
  }
  return NULL;
}

int orc_knn_bruteforce(const float *X, uint64_t N, int L, const float *Q, uint32_t nq,
                       uint32_t k, uint64_t *ids, float *d2, int nthreads)
{
  if (k == 0 || L % 8 != 0) return -1;
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > N) nthreads = N ? (int)N : 1;
  struct bf_job *jobs = calloc(nthreads, sizeof(struct bf_job));
  pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
  uint64_t per = (N + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    uint64_t b = (uint64_t)t * per, e = b + per;
    if (b > N) b = N;
    if (e > N) e = N;
    jobs[t] = (struct bf_job){X, Q, b, e, L, nq, k,
                              malloc(sizeof(float) * nq * k), malloc(sizeof(uint64_t) * nq * k),
                              calloc(nq, sizeof(uint32_t))};
    pthread_create(&th[t], NULL, orc_bf_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  /* merge: chunks are in ascending id order, so visiting thread lists in
     order and inserting with the same rule keeps the smallest id */
  for (uint32_t qi = 0; qi < nq; ++qi) {
    float *dd = d2 + (uint64_t)qi * k;
    uint64_t *ii = ids + (uint64_t)qi * k;
    uint32_t cur = 0;
    for (int t = 0; t < nthreads; ++t) {
      struct bf_job *j = &jobs[t];
      for (uint32_t r = 0; r < j->cur[qi]; ++r) {
        float nd = j->d[(uint64_t)qi * k + r];
        uint64_t nid = j->id[(uint64_t)qi * k + r];
        float kth = cur == k ? dd[k - 1] : FLT_MAX;
        if (nd < kth) orc_topk_insert(dd, ii, &cur, k, nd, nid);
      }
    }
    for (uint32_t r = cur; r < k; ++r) {
      dd[r] = FLT_MAX;
      ii[r] = UINT64_MAX;
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    free(jobs[t].d);
    free(jobs[t].id);
    free(jobs[t].cur);
  }
  free(jobs);
  free(th);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* approximate query modes (epsilon, ng) on an index tree                    */
/* ------------------------------------------------------------------------ */

/* pqueue.c:63-200 with cmp_pri = (next > curr) (hercules_query_engine.h:36-39):
   a binary min-heap in d[1..size-1] ordered by the bound only, so items of
   equal bound leave in the order the heap's moves give them */
struct orc_pq_item {
  float pri;
  int32_t node;
};
struct orc_pq {
  struct orc_pq_item *d;
  size_t size; /* 1 + number of items */
};

static void orc_pq_bubble_up(struct orc_pq *q, size_t i)
{
  struct orc_pq_item mv = q->d[i];
  while (i > 1 && (double)q->d[i >> 1].pri > (double)mv.pri) {
    q->d[i] = q->d[i >> 1];
    i >>= 1;
  }
  q->d[i] = mv;
}

static size_t orc_pq_min_child(const struct orc_pq *q, size_t i)
{
  size_t c = i << 1;
  if (c >= q->size) return 0;
  if (c + 1 < q->size && (double)q->d[c].pri > (double)q->d[c + 1].pri) ++c;
  return c;
}

static void orc_pq_percolate_down(struct orc_pq *q, size_t i)
{
  struct orc_pq_item mv = q->d[i];
  size_t c;
  while ((c = orc_pq_min_child(q, i)) && (double)mv.pri > (double)q->d[c].pri) {
    q->d[i] = q->d[c];
    i = c;
  }
  q->d[i] = mv;
}

static void orc_pq_insert(struct orc_pq *q, float pri, int32_t node)
{
  size_t i = q->size++;
  q->d[i].pri = pri;
  q->d[i].node = node;
  orc_pq_bubble_up(q, i);
}

static int orc_pq_pop(struct orc_pq *q, struct orc_pq_item *out)
{
  if (q->size == 1) return 0;
  *out = q->d[1];
  q->d[1] = q->d[--q->size];
  orc_pq_percolate_down(q, 1);
  return 1;
}

enum { OF_LEAF = 0, OF_LEFT, OF_RIGHT, OF_SEGOFF, OF_NSEG, OF_FROM, OF_TO, OF_IDX, OF_POS, OF_SIZE };

struct orc_tree {
  const int64_t *nd;
  const float *split_value;
  const int16_t *seg_end;
  const float *ind;
  const float *X;
  const uint8_t *sax;
  int L;
  uint32_t nnodes;
  uint32_t nleaves;    /* leaves with rows (the reference's leaves_counter) */
  uint64_t nseries;    /* rows under the root */
  float bp[255];
};

static const int64_t *orc_node(const struct orc_tree *t, int32_t i) { return t->nd + (int64_t)i * ORC_NODE_FIELDS; }

/* calculate_node_min_distance (hercules_node.c:879-924) in the FMA build */
static float orc_node_lb(const struct orc_tree *t, int32_t i, const float *q)
{
  const int64_t *n = orc_node(t, i);
  if (n[OF_NSEG] == 0) return 0.0f; /* empty leaf: no synopsis on disk */
  return orc_lb_eapca(q, (int)n[OF_NSEG], t->seg_end + n[OF_SEGOFF], t->ind + 4 * n[OF_SEGOFF], 1);
}

/* the running answer of a search: distinct distances, FLT_MAX past the
   filled ranks (queue_bounded_sorted_insert, hercules_query_engine.c:4034-4082) */
struct orc_run {
  float *d;
  uint64_t *id;
  uint32_t cur, k;
};

static float orc_kth(const struct orc_run *r) { return r->cur == r->k ? r->d[r->k - 1] : FLT_MAX; }

/* calculate_node_knn_distance (hercules_node.c:708-790), the leaf loop of
   approximate_greedy_knn_search (hercules_query_engine.c:111-139) and of
   exact_de_knn_search_noapprox (:2484-2500): every row against the k-th
   distance of the moment, inserted when below it.  Returns the k-th
   distance read before the last row (the value the caller's variable
   keeps, :2486) or `kth_in` for an empty leaf. */
static float orc_scan_leaf(const struct orc_tree *t, int32_t i, const float *q, struct orc_run *r, float kth_in)
{
  const int64_t *n = orc_node(t, i);
  float bound = kth_in;
  for (int64_t j = 0; j < n[OF_SIZE]; ++j) {
    uint64_t row = (uint64_t)n[OF_POS] + (uint64_t)j;
    bound = orc_kth(r);
    float d = orc_ed2(q, t->X + row * (uint64_t)t->L, t->L, bound);
    if (d < bound) orc_topk_insert(r->d, r->id, &r->cur, r->k, d, row);
  }
  return bound;
}

/* node_split_policy_route_to_left (hercules_node_split.c:172-194) */
static int orc_route_left(const struct orc_tree *t, int32_t i, const float *q)
{
  const int64_t *n = orc_node(t, i);
  float ind[2];
  orc_seg_mean_std(q, (int)n[OF_FROM], (int)n[OF_TO], &ind[0], &ind[1]);
  return ind[n[OF_IDX] & 1] < t->split_value[i];
}

struct orc_mode_params {
  uint32_t k, nprobes;
  float epsilon, eapca_th, sax_th;
  int approx_budget;
};

struct orc_cand {
  int32_t node;
  float lb;
};

static int orc_cand_by_pos(const void *a, const void *b, void *tv)
{
  const struct orc_tree *t = (const struct orc_tree *)tv;
  int64_t pa = orc_node(t, ((const struct orc_cand *)a)->node)[OF_POS];
  int64_t pb = orc_node(t, ((const struct orc_cand *)b)->node)[OF_POS];
  return pa < pb ? -1 : (pa > pb ? 1 : 0);
}

/* one query of a mode: marks the leaves whose rows the reference's search
   examines, then answers over exactly those rows.  info[4]: leaves scanned
   before the candidate stage, candidate leaves, the branch the candidates
   took (0 none, 1 sequential, 2 iSAX + parallel refinement), candidate
   leaves scanned. */
static void orc_mode_query(const struct orc_tree *t, const float *q, const struct orc_mode_params *p,
                           uint8_t *mark, struct orc_pq *pq, struct orc_cand *cand, uint64_t *ids, float *d2,
                           uint32_t *info)
{
  const uint32_t k = p->k;
  struct orc_run run = {malloc(sizeof(float) * k), malloc(sizeof(uint64_t) * k), 0, k};
  memset(mark, 0, t->nnodes);
  memset(info, 0, 4 * sizeof(uint32_t));
  pq->size = 1;
  uint32_t nvis = 0, ncand = 0;
  struct orc_pq_item it;
  if (p->nprobes == 0) {
    /* approximate_greedy_knn_search (hercules_query_engine.c:70-174, with the
       count kept in the caller's variable): best-first from the root, every
       row of each popped leaf, up to approx_budget leaves */
    orc_pq_insert(pq, orc_node_lb(t, 0, q), 0);
    while (orc_pq_pop(pq, &it)) {
      if (it.pri > orc_kth(&run)) break; /* the popped item is dropped */
      const int64_t *n = orc_node(t, it.node);
      if (n[OF_LEAF]) {
        (void)orc_scan_leaf(t, it.node, q, &run, FLT_MAX);
        mark[it.node] = 1;
        ++nvis;
        if ((int)nvis >= p->approx_budget) break;
      } else {
        float kth = orc_kth(&run);
        for (int c = 0; c < 2; ++c) {
          int32_t ch = (int32_t)n[OF_LEFT + c];
          float lb = orc_node_lb(t, ch, q);
          if (lb < kth) orc_pq_insert(pq, lb, ch);
        }
      }
    }
    /* exact_de_knn_search_psq_lf_sims_adaptive (:836-891): the same queue,
       bounds against kth/(1+epsilon) of the approximate answer; the leaves
       that pass are the candidates.  A leaf scanned above is never pushed
       again (a node has one parent, expanded once), so the test against the
       approximate leaf (:866, :878) changes nothing. */
    const float kth0 = orc_kth(&run);
    const float thr = kth0 / (1.0f + p->epsilon);
    while (orc_pq_pop(pq, &it)) {
      if (it.pri > thr) break;
      const int64_t *n = orc_node(t, it.node);
      if (n[OF_LEAF]) {
        cand[ncand].node = it.node;
        cand[ncand].lb = it.pri;
        ++ncand;
      } else {
        for (int c = 0; c < 2; ++c) {
          int32_t ch = (int32_t)n[OF_LEFT + c];
          float lb = orc_node_lb(t, ch, q);
          if (lb < thr) orc_pq_insert(pq, lb, ch);
        }
      }
    }
    info[1] = ncand;
    if (ncand) {
      /* :893-921: sequential when the leaf pruning is weak, else the iSAX
         pass counts the rows whose bound is within kth0 (:2171-2205) and
         goes sequential when that pruning is weak (:967-978) */
      qsort_r(cand, ncand, sizeof(struct orc_cand), orc_cand_by_pos, (void *)t);
      float flp = (float)(1.0 - (double)ncand / (double)(float)t->nleaves);
      int sequential = flp < p->eapca_th;
      if (!sequential) {
        float paa[16];
        orc_paa(q, t->L, 16, paa);
        uint64_t lab = 0;
        for (uint32_t c = 0; c < ncand; ++c) {
          const int64_t *n = orc_node(t, cand[c].node);
          if (!(cand[c].lb <= kth0)) continue;
          for (int64_t j = 0; j < n[OF_SIZE]; ++j) {
            float lbs = orc_lb_sax(paa, t->sax + ((uint64_t)n[OF_POS] + (uint64_t)j) * 16, t->L, t->bp);
            if (lbs <= kth0) ++lab;
          }
        }
        float slp = (float)(1.0 - (double)lab / (double)(float)t->nseries);
        sequential = slp < p->sax_th;
      }
      info[2] = sequential ? 1 : 2;
      if (sequential) {
        /* exact_de_knn_search_noapprox (:2450-2507): candidates in file
           order, each tested against the k-th distance its loop variable
           holds, which lags the last insertion of the previous leaf */
        float kth_b = orc_kth(&run);
        for (uint32_t c = 0; c < ncand; ++c) {
          if (cand[c].lb <= kth_b / (1.0f + p->epsilon)) {
            kth_b = orc_scan_leaf(t, cand[c].node, q, &run, kth_b);
            mark[cand[c].node] = 1;
            ++info[3];
          }
        }
      } else {
        /* the iSAX and refinement workers (:1960-2030, :2140-2213) prune
           with exact bounds only: every candidate's rows count */
        for (uint32_t c = 0; c < ncand; ++c) mark[cand[c].node] = 1;
        info[3] = ncand;
      }
    }
  } else {
    /* exact_ng_knn_search (hercules_query_engine.c:2623-2799): the leaf the
       split policies route the query to (approximate_knn_search, :176-223),
       then best-first from the root until nprobes leaves have been scanned,
       stopping at a bound above the first leaf's best distance */
    int32_t i = 0;
    while (!orc_node(t, i)[OF_LEAF]) i = (int32_t)orc_node(t, i)[orc_route_left(t, i, q) ? OF_LEFT : OF_RIGHT];
    const int32_t first = i;
    (void)orc_scan_leaf(t, first, q, &run, FLT_MAX);
    mark[first] = 1;
    nvis = 1;
    const float best = run.cur ? run.d[0] : FLT_MAX;
    const int32_t best_leaf = run.cur ? first : -1;
    orc_pq_insert(pq, orc_node_lb(t, 0, q), 0);
    while (orc_pq_pop(pq, &it) && nvis < p->nprobes) {
      if (it.pri > best) break;
      const int64_t *n = orc_node(t, it.node);
      if (n[OF_LEAF]) {
        (void)orc_scan_leaf(t, it.node, q, &run, FLT_MAX);
        mark[it.node] = 1;
        ++nvis;
      } else {
        float kth = orc_kth(&run);
        for (int c = 0; c < 2; ++c) {
          int32_t ch = (int32_t)n[OF_LEFT + c];
          float lb = orc_node_lb(t, ch, q);
          if (lb < kth && ch != best_leaf) orc_pq_insert(pq, lb, ch);
        }
      }
    }
  }
  info[0] = nvis;
  /* the answer: the k smallest distinct distances over the marked leaves'
     rows (the streaming inserts above keep exactly those), each with the
     smallest row of that distance: marked leaves again, rows ascending */
  run.cur = 0;
  uint32_t nm = 0;
  for (uint32_t j = 0; j < t->nnodes; ++j)
    if (mark[j]) {
      cand[nm].node = (int32_t)j;
      cand[nm].lb = 0.0f;
      ++nm;
    }
  qsort_r(cand, nm, sizeof(struct orc_cand), orc_cand_by_pos, (void *)t);
  for (uint32_t a = 0; a < nm; ++a) (void)orc_scan_leaf(t, cand[a].node, q, &run, FLT_MAX);
  for (uint32_t r = 0; r < k; ++r) {
    d2[r] = r < run.cur ? run.d[r] : FLT_MAX;
    ids[r] = r < run.cur ? run.id[r] : UINT64_MAX;
  }
  free(run.d);
  free(run.id);
}

struct mode_job {
  const struct orc_tree *t;
  const struct orc_mode_params *p;
  uint32_t q0, q1;
  const float *Q;
  uint64_t *ids;
  float *d2;
  uint32_t *info;
};

static void *orc_mode_worker(void *arg)
{
  struct mode_job *j = (struct mode_job *)arg;
  const uint32_t nn = j->t->nnodes;
  uint8_t *mark = malloc(nn);
  struct orc_pq pq = {malloc(sizeof(struct orc_pq_item) * (nn + 2)), 1};
  struct orc_cand *cand = malloc(sizeof(struct orc_cand) * (nn + 1));
  const uint32_t k = j->p->k;
  for (uint32_t qi = j->q0; qi < j->q1; ++qi)
    orc_mode_query(j->t, j->Q + (uint64_t)qi * j->t->L, j->p, mark, &pq, cand, j->ids + (uint64_t)qi * k,
                   j->d2 + (uint64_t)qi * k, j->info + 4 * (uint64_t)qi);
  free(mark);
  free(pq.d);
  free(cand);
  return NULL;
}

int orc_knn_mode(const int64_t *nodes, const float *split_value, uint32_t nnodes, const int16_t *seg_end,
                 const float *ind, const float *X, const uint8_t *sax, int L, const float *Q, uint32_t nq,
                 uint32_t k, float epsilon, int approx_budget, uint32_t nprobes, float eapca_th, float sax_th,
                 uint64_t *ids, float *d2, uint32_t *info, int nthreads)
{
  if (k == 0 || L % 8 != 0 || nnodes == 0 || !(epsilon >= 0.0f)) return -1;
  if (nthreads < 1) nthreads = 1;
  if ((uint32_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
  struct orc_tree t;
  memset(&t, 0, sizeof(t));
  t.nd = nodes;
  t.split_value = split_value;
  t.seg_end = seg_end;
  t.ind = ind;
  t.X = X;
  t.sax = sax;
  t.L = L;
  t.nnodes = nnodes;
  t.nseries = (uint64_t)nodes[OF_SIZE];
  for (uint32_t i = 0; i < nnodes; ++i)
    if (nodes[(int64_t)i * ORC_NODE_FIELDS + OF_LEAF] && nodes[(int64_t)i * ORC_NODE_FIELDS + OF_NSEG]) ++t.nleaves;
  orc_breakpoints256(t.bp);
  struct orc_mode_params p = {k, nprobes, epsilon, eapca_th, sax_th, approx_budget};
  struct mode_job *jobs = calloc(nthreads, sizeof(struct mode_job));
  pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
  uint32_t per = (nq + nthreads - 1) / nthreads;
  for (int i = 0; i < nthreads; ++i) {
    uint32_t a = i * per < nq ? i * per : nq, b = a + per < nq ? a + per : nq;
    jobs[i] = (struct mode_job){&t, &p, a, b, Q, ids, d2, info};
    pthread_create(&th[i], NULL, orc_mode_worker, &jobs[i]);
  }
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  free(jobs);
  free(th);
  return 0;
}
