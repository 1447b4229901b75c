/* msv_oracle.c -- scalar CPU restatement of HMMER3's p7_MSVFilter composite.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; the product never links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * It restates, one cell at a time, what the SSE reference computes sixteen
 * lanes at a time:
 *
 *   orc_ssv        p7_SSVFilter          impl_sse/ssvfilter.c:1364-1413
 *                  (cell step STEP_SINGLE :445-447; shift-with-OR at the
 *                  stripe boundary CONVERT_STEP :590-596; sbv construction
 *                  sf_conversion, impl_sse/p7_oprofile.c:575-587)
 *   orc_msv_bytes  byte MSV row loop      impl_sse/msvfilter.c:112-210
 *   orc_msv        the composite          impl_sse/msvfilter.c:103-108
 *   orc_tjb        ReconfigMSVLength      impl_sse/p7_oprofile.c:921-925
 *                  via unbiased_byteify   impl_sse/p7_oprofile.c:512-519
 *   orc_score      score expression       msvfilter.c:208-210 (= ssvfilter.c:1408-1410)
 *   orc_nullsc     p7_bg_SetLength + p7_bg_NullOne   p7_bg.c:191-198, 359-363
 *   orc_pass       MSV stage of p7_Pipeline          p7_pipeline.c:667-677,
 *                  esl_gumbel_surv                   easel/esl_gumbel.c:129-137
 *
 * Striped layout: the reference keeps Q = max(2, ceil(M/16)) vectors of 16
 * bytes (impl_sse.h:26); byte z of vector q is model position
 * k = q + z*Q + 1, and positions M+1..16Q are padding with cost 255
 * (p7_oprofile.c:628).  Both scalar loops below run over all 16Q positions so
 * the padding behaves exactly as in the SIMD code.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "msv_oracle.h"

static int nqb(int M) { int q = (M - 1) / 16 + 1; return q < 2 ? 2 : q; }

static inline uint8_t sat_addu8(int a, int b) { int s = a + b; return (uint8_t) (s > 255 ? 255 : s); }
static inline uint8_t sat_subu8(int a, int b) { int s = a - b; return (uint8_t) (s < 0 ? 0 : s); }
static inline int8_t  sat_subi8(int a, int b) { int s = a - b; return (int8_t) (s > 127 ? 127 : (s < -128 ? -128 : s)); }

/* rbv cost for residue x at model position k (1-based), 255 in padding. */
static inline int cost(const orc_profile *p, int x, int k) { return k <= p->M ? p->rbv[x * p->M + k - 1] : 255; }

/* unbiased_byteify(om, logf(3/(L+3))), p7_oprofile.c:512-519 / :923. */
int orc_tjb(int L)
{
  float scale_b = (float) (3.0 / 0.69314718055994529);
  float sc = logf(3.0f / (float) (L + 3));
  sc = -1.0f * roundf(scale_b * sc);
  return (sc > 255.) ? 255 : (uint8_t) sc;
}

/* msvfilter.c:208-210. */
float orc_score(int xJ, int tjb, int base_b, float scale_b)
{
  float sc = ((float) (xJ - tjb) - (float) base_b);
  sc /= scale_b;
  sc -= 3.0;
  return sc;
}

/* p7_SSVFilter, ssvfilter.c:1364-1413, as a scalar diagonal DP.
 * sbv = ((127 + bias) -sat rbv) ^ 127 read as int8 (p7_oprofile.c:583).
 * Each cell: s = sat_i8(s_in - sbv); xE = max_u8(xE, (uint8) s).  Every
 * diagonal starts at beginv = 0x80 in row 1; a value crossing from the last
 * vector of a stripe into the next stripe's first vector is shifted and OR'ed
 * with 0x80 (CONVERT_STEP), and lane 0 of vector 0 takes 0x80. */
int orc_ssv(const orc_profile *p, const uint8_t *x, int L, float *usc)
{
  int M = p->M, Q = nqb(M), Mp = 16 * Q;
  int tjb = orc_tjb(L);
  int i, k;
  uint16_t xE, xJ;
  uint8_t u = 128;
  int8_t *prev, *cur, *tmp;
  int8_t sbv;

  if (p->tbm_b + tjb + p->tec_b + p->bias_b >= 127) return ORC_eslENORESULT;   /* :1370 */

  prev = malloc(Mp + 1);
  cur  = malloc(Mp + 1);
  for (k = 0; k <= Mp; k++) prev[k] = (int8_t) 0x80;
  for (i = 0; i < L; i++) {
    int xi = x[i];
    for (k = 1; k <= Mp; k++) {
      int s_in;
      if (k == 1) s_in = (int8_t) 0x80;
      else {
        s_in = prev[k - 1];
        if ((k - 1) % Q == 0) s_in = (int8_t) ((uint8_t) s_in | 0x80);
      }
      sbv = (int8_t) (sat_subu8(127 + p->bias_b, cost(p, xi, k)) ^ 127);
      cur[k] = sat_subi8(s_in, sbv);
      if ((uint8_t) cur[k] > u) u = (uint8_t) cur[k];
    }
    tmp = prev; prev = cur; cur = tmp;
  }
  free(prev); free(cur);

  xE = u;
  if (xE >= 255 - p->bias_b) {                                               /* :1379 */
    *usc = INFINITY;
    if (p->base_b - tjb - p->tbm_b < 128) return ORC_eslENORESULT;            /* :1383 */
    return ORC_eslERANGE;
  }
  xE += p->base_b - tjb - p->tbm_b;                                          /* :1393 */
  xE -= 128;
  if (xE >= 255 - p->bias_b) { *usc = INFINITY; return ORC_eslERANGE; }      /* :1396 */
  xJ = xE - p->tec_b;
  if (xJ > p->base_b) return ORC_eslENORESULT;                               /* :1405 */
  *usc = orc_score(xJ, tjb, p->base_b, p->scale_b);
  return ORC_eslOK;
}

/* The byte MSV loop of p7_MSVFilter, msvfilter.c:112-210.  Cell (i,k) takes
 * the previous row's k-1 (0 at k = 1, the zero shifted in by
 * _mm_slli_si128 at :146). */
int orc_msv_bytes(const orc_profile *p, const uint8_t *x, int L, float *usc)
{
  int M = p->M, Q = nqb(M), Mp = 16 * Q;
  int tjb = orc_tjb(L);
  int tjbm = (uint8_t) (tjb + p->tbm_b);                 /* int8 sum splat, :120 */
  uint8_t xJ = 0, xB, xE;
  uint8_t *dp = calloc(Mp + 1, 1);
  int i, k;

  xB = sat_subu8(p->base_b, tjbm);                                            /* :124 */
  for (i = 0; i < L; i++) {
    int xi = x[i];
    uint8_t mpv = 0;                /* previous row, position k-1 */
    xE = 0;
    for (k = 1; k <= Mp; k++) {
      uint8_t sv = mpv > xB ? mpv : xB;                                        /* :150 */
      sv = sat_addu8(sv, p->bias_b);                                          /* :151 */
      sv = sat_subu8(sv, cost(p, xi, k));                                     /* :152 */
      if (sv > xE) xE = sv;                                                   /* :153 */
      mpv = dp[k];
      dp[k] = sv;
    }
    if (sat_addu8(xE, p->bias_b) == 255) {                                    /* :160-185 */
      free(dp);
      *usc = INFINITY;
      return ORC_eslERANGE;
    }
    xE = sat_subu8(xE, p->tec_b);                                             /* :187 */
    if (xE > xJ) xJ = xE;                                                     /* :188 */
    xB = sat_subu8(p->base_b > xJ ? p->base_b : xJ, tjbm);                    /* :190-191 */
  }
  free(dp);
  *usc = orc_score(xJ, tjb, p->base_b, p->scale_b);
  return ORC_eslOK;
}

/* p7_MSVFilter: SSV first, byte MSV when SSV has no result (msvfilter.c:107-108). */
int orc_msv(const orc_profile *p, const uint8_t *x, int L, float *usc)
{
  int status = orc_ssv(p, x, L, usc);
  if (status != ORC_eslENORESULT) return status;
  return orc_msv_bytes(p, x, L, usc);
}

/* p7_bg_SetLength (p7_bg.c:191-198) then p7_bg_NullOne (:359-363). */
float orc_nullsc(int L)
{
  float p1 = (float) L / (float) (L + 1);
  return (float) L * log(p1) + log(1. - p1);
}

/* esl_gumbel_surv, easel/esl_gumbel.c:129-137. */
static double gumbel_surv(double x, double mu, double lambda)
{
  double y  = lambda * (x - mu);
  double ey = -exp(-y);
  if (fabs(ey) < 5e-9) return -ey;
  else                 return 1 - exp(ey);
}

/* p7_pipeline.c:667-677: pass unless P > F1. */
int orc_pass(float usc, int L, float mu, float lambda, double F1)
{
  float nullsc = orc_nullsc(L);
  float seq_score = (usc - nullsc) / 0.69314718055994529;
  double P = gumbel_surv(seq_score, mu, lambda);
  return (P > F1) ? 0 : 1;
}

/* ---- threaded batch over many targets ---- */
typedef struct {
  const orc_profile *p;
  const uint8_t *res; const int64_t *off; const int32_t *len;
  float *usc; int32_t *status;
  int n, next;
  pthread_mutex_t lock;
} orc_batch;

static void *orc_worker(void *arg)
{
  orc_batch *b = arg;
  for (;;) {
    int i0, i1, i;
    pthread_mutex_lock(&b->lock);
    i0 = b->next; b->next += 64; if (b->next > b->n) b->next = b->n; i1 = b->next;
    pthread_mutex_unlock(&b->lock);
    if (i0 >= i1) break;
    for (i = i0; i < i1; i++)
      b->status[i] = orc_msv(b->p, b->res + b->off[i], b->len[i], &b->usc[i]);
  }
  return NULL;
}

int orc_msv_batch(const orc_profile *p, const uint8_t *res, const int64_t *off,
                  const int32_t *len, int n, float *usc, int32_t *status, int nthreads)
{
  orc_batch b;
  pthread_t *tid;
  int t;
  if (nthreads < 1) nthreads = 1;
  b.p = p; b.res = res; b.off = off; b.len = len; b.usc = usc; b.status = status;
  b.n = n; b.next = 0;
  pthread_mutex_init(&b.lock, NULL);
  tid = malloc(sizeof(pthread_t) * nthreads);
  for (t = 0; t < nthreads; t++) pthread_create(&tid[t], NULL, orc_worker, &b);
  for (t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
  free(tid);
  pthread_mutex_destroy(&b.lock);
  return 0;
}

/* ---- checker for the GPU decomposition (not a reference restatement) ----
 * The HIP path computes, per pair, Rp = max over all cells of
 * W(i,k) - B0 with W(i,k) = max(W(i-1,k-1) + bias - rbv[x_i][k], B0) and
 * B0 = max(0, base - ((tjb + tbm) mod 256)), then reaches the composite's
 * result from Rp alone unless the J state may fire.  These two functions
 * state that in plain integers so tests can confirm on the CPU that
 * decide(jfree(x)) equals orc_msv(x) whenever it does not defer. */
int orc_jfree_max(const orc_profile *p, const uint8_t *x, int L)
{
  int M = p->M, i, k;
  int tjb = orc_tjb(L);
  int B0 = p->base_b - ((tjb + p->tbm_b) & 255);
  int *prev = malloc(sizeof(int) * (M + 1)), *cur = malloc(sizeof(int) * (M + 1)), *t;
  int best = 0;
  if (B0 < 0) B0 = 0;
  for (k = 0; k <= M; k++) prev[k] = B0;
  for (i = 0; i < L; i++) {
    cur[0] = B0;
    for (k = 1; k <= M; k++) {
      int w = prev[k - 1] + p->bias_b - p->rbv[x[i] * M + k - 1];
      if (w < B0) w = B0;
      cur[k] = w;
      if (w - B0 > best) best = w - B0;
    }
    t = prev; prev = cur; cur = t;
  }
  free(prev); free(cur);
  return best;
}

/* Returns 0 (eslOK, *xJ set), 16 (eslERANGE) or 19 (defer to the byte MSV). */
int orc_decide(const orc_profile *p, int Rp, int L, int *xJ)
{
  int tjb = orc_tjb(L);
  int tjbm = (tjb + p->tbm_b) & 255;
  int B0 = p->base_b - tjbm;
  int R, guard;
  if (B0 < 0) B0 = 0;
  R = Rp + B0;
  guard = (tjb + p->tbm_b + p->tec_b + p->bias_b) >= 127;
  if (R >= 255 - p->bias_b) return ORC_eslERANGE;
  if (R - p->tec_b > p->base_b) return ORC_eslENORESULT;
  if (guard) {
    if (Rp == 0) return ORC_eslENORESULT;
    *xJ = (R - p->tec_b) > 0 ? R - p->tec_b : 0;
    return ORC_eslOK;
  }
  *xJ = R - p->tec_b;
  return ORC_eslOK;
}
