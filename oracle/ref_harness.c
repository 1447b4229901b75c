/* ref_harness.c -- a thin C wrapper over the compiled HMMER 3.1b1 reference.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is linked against objects compiled
 * from the reference sources where they lie (see oracle/Makefile) into
 * oracle/_ref/libmsvref.so.  It is used as a checker (golden-vector
 * generation, parity tests) and as the CPU baseline in bench.py.  Nothing in
 * the product links or loads it.
 *
 * Every entry point only calls reference API functions; none of the
 * reference's algorithms is reimplemented here.
 *
 *   ref_profile_sample   p7_hmm_Sample (p7_hmm.c:877) -> p7_Calibrate
 *                        (evalues.c:60) -> p7_ProfileConfig (modelconfig.c:49)
 *                        -> p7_oprofile_Convert (impl_sse/p7_oprofile.c:840)
 *   ref_msv              p7_oprofile_ReconfigLength (p7_oprofile.c:897) +
 *                        p7_MSVFilter (impl_sse/msvfilter.c:78)
 *   ref_ssv              p7_SSVFilter (impl_sse/ssvfilter.c:1364)
 *   ref_pass             MSV stage of p7_Pipeline (p7_pipeline.c:662-677)
 *   ref_msv_batch        p7_MSVFilter over many targets on N pthreads, each
 *                        with its own p7_oprofile_Clone / p7_omx (the way
 *                        hmmsearch.c:429-439 gives each worker its own copy)
 *   ref_iid / ref_core_emit  esl_rsq_xfIID (esl_randomseq.c:764) and
 *                        p7_CoreEmit (emit.c:71) for generator pinning
 *   ref_vit              p7_oprofile_ReconfigLength + p7_ViterbiFilter
 *                        (impl_sse/vitfilter.c:84); ref_vit_batch on N threads
 *   ref_vit_fields       the ViterbiFilter part of P7_OPROFILE (rwv, twv, xw,
 *                        scale_w, base_w, ddbound_w, nj), de-striped
 *   ref_profile_tsc      the generic profile's transition scores (gm->tsc)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>

#include "easel.h"
#include "esl_alphabet.h"
#include "esl_gumbel.h"
#include "esl_random.h"
#include "esl_randomseq.h"
#include "esl_sq.h"
#include "hmmer.h"

typedef struct {
  ESL_ALPHABET *abc;
  P7_HMM       *hmm;
  P7_BG        *bg;
  P7_PROFILE   *gm;
  P7_OPROFILE  *om;
  P7_OMX       *ox;
} REFPROF;

static void refprof_free(REFPROF *h)
{
  if (!h) return;
  if (h->ox)  p7_omx_Destroy(h->ox);
  if (h->om)  p7_oprofile_Destroy(h->om);
  if (h->gm)  p7_profile_Destroy(h->gm);
  if (h->bg)  p7_bg_Destroy(h->bg);
  if (h->hmm) p7_hmm_Destroy(h->hmm);
  if (h->abc) esl_alphabet_Destroy(h->abc);
  free(h);
}

/* Configure gm/om for an HMM that is already in h->hmm. */
static int refprof_configure(REFPROF *h, int config_L, int calibrate)
{
  if (calibrate) {
    if (p7_Calibrate(h->hmm, NULL, NULL, NULL, NULL, NULL) != eslOK) return -1;
  }
  h->bg = p7_bg_Create(h->abc);
  h->gm = p7_profile_Create(h->hmm->M, h->abc);
  if (p7_ProfileConfig(h->hmm, h->bg, h->gm, config_L, p7_LOCAL) != eslOK) return -1;
  h->om = p7_oprofile_Create(h->hmm->M, h->abc);
  if (p7_oprofile_Convert(h->gm, h->om) != eslOK) return -1;
  h->ox = p7_omx_Create(h->hmm->M, 0, 0);
  return 0;
}

/* Sample an amino-acid HMM of length M from esl_randomness_Create(seed). */
void *ref_profile_sample(uint32_t seed, int M, int config_L, int calibrate)
{
  REFPROF *h = calloc(1, sizeof(REFPROF));
  ESL_RANDOMNESS *r = esl_randomness_Create(seed);
  h->abc = esl_alphabet_Create(eslAMINO);
  if (p7_hmm_Sample(r, M, h->abc, &h->hmm) != eslOK) { esl_randomness_Destroy(r); refprof_free(h); return NULL; }
  esl_randomness_Destroy(r);
  if (refprof_configure(h, config_L, calibrate) != 0) { refprof_free(h); return NULL; }
  return h;
}

/* Build a profile from caller-given HMM probabilities (mat: (M+1) x 20,
 * row 0 unused; t: (M+1) x 7 in p7H_{MM,MI,MD,IM,II,DM,DD} order).  Used to
 * pin the conversion on hand-made edge-case models. */
void *ref_profile_from_probs(int M, const float *mat, const float *t, int config_L, int calibrate)
{
  REFPROF *h = calloc(1, sizeof(REFPROF));
  int k, x;
  h->abc = esl_alphabet_Create(eslAMINO);
  h->hmm = p7_hmm_Create(M, h->abc);
  for (k = 0; k <= M; k++) {
    for (x = 0; x < 20; x++) {
      h->hmm->mat[k][x] = (k == 0) ? (x == 0 ? 1.0f : 0.0f) : mat[k * 20 + x];
      h->hmm->ins[k][x] = 0.05f;
    }
    for (x = 0; x < 7; x++) h->hmm->t[k][x] = t[k * 7 + x];
  }
  p7_hmm_SetName(h->hmm, "edge-hmm");
  p7_hmm_SetConsensus(h->hmm, NULL);
  if (refprof_configure(h, config_L, calibrate) != 0) { refprof_free(h); return NULL; }
  return h;
}

void ref_profile_free(void *vh) { refprof_free((REFPROF *) vh); }

int ref_profile_M(void *vh) { return ((REFPROF *) vh)->om->M; }

/* rbv de-striped to row-major out[x * M + (k-1)], x in 0..Kp-1 (29 rows). */
void ref_profile_rbv(void *vh, uint8_t *out)
{
  REFPROF *h = vh;
  P7_OPROFILE *om = h->om;
  int M = om->M, Q = p7O_NQB(M), Kp = om->abc->Kp;
  int x, q, z;
  for (x = 0; x < Kp; x++)
    for (q = 0; q < Q; q++) {
      union { __m128i v; uint8_t b[16]; } u;
      u.v = om->rbv[x][q];
      for (z = 0; z < 16; z++) {
        int k = q + z * Q + 1;
        if (k <= M) out[x * M + (k - 1)] = u.b[z];
      }
    }
}

/* Scalars: out_i = {tbm_b, tec_b, tjb_b, base_b, bias_b}; out_f = {scale_b,
 * mmu, mlambda}.  tjb_b is for the L the profile was last configured with. */
void ref_profile_scalars(void *vh, int *out_i, float *out_f)
{
  P7_OPROFILE *om = ((REFPROF *) vh)->om;
  out_i[0] = om->tbm_b; out_i[1] = om->tec_b; out_i[2] = om->tjb_b;
  out_i[3] = om->base_b; out_i[4] = om->bias_b;
  out_f[0] = om->scale_b; out_f[1] = om->evparam[p7_MMU]; out_f[2] = om->evparam[p7_MLAMBDA];
}

/* HMM match emissions, out[(k) * 20 + x] for k = 1..M (row 0 zeroed), and
 * transitions t[k * 7 + j] for k = 0..M. */
void ref_profile_hmm(void *vh, float *mat, float *t)
{
  P7_HMM *hmm = ((REFPROF *) vh)->hmm;
  int k, x;
  for (k = 0; k <= hmm->M; k++) {
    for (x = 0; x < 20; x++) mat[k * 20 + x] = (k == 0) ? 0.0f : hmm->mat[k][x];
    for (x = 0; x < 7; x++)  t[k * 7 + x] = hmm->t[k][x];
  }
}

/* Generic-profile match scores p7P_MSC, out[x * M + (k-1)], x in 0..Kp-1. */
void ref_profile_msc(void *vh, float *out)
{
  P7_PROFILE *gm = ((REFPROF *) vh)->gm;
  int k, x;
  for (x = 0; x < gm->abc->Kp; x++)
    for (k = 1; k <= gm->M; k++) out[x * gm->M + (k - 1)] = p7P_MSC(gm, k, x);
}

/* tjb_b after p7_oprofile_ReconfigLength(om, L). */
int ref_tjb(void *vh, int L)
{
  P7_OPROFILE *om = ((REFPROF *) vh)->om;
  p7_oprofile_ReconfigLength(om, L);
  return om->tjb_b;
}

/* p7_MSVFilter on dsq[0..L+1] (digital, sentinels at 0 and L+1). */
int ref_msv(void *vh, const uint8_t *dsq, int L, float *usc)
{
  REFPROF *h = vh;
  p7_oprofile_ReconfigLength(h->om, L);
  return p7_MSVFilter((const ESL_DSQ *) dsq, L, h->om, h->ox, usc);
}

/* p7_SSVFilter alone: returns eslOK / eslERANGE / eslENORESULT. */
int ref_ssv(void *vh, const uint8_t *dsq, int L, float *usc)
{
  REFPROF *h = vh;
  p7_oprofile_ReconfigLength(h->om, L);
  *usc = 0.0f;
  return p7_SSVFilter((const ESL_DSQ *) dsq, L, h->om, usc);
}

/* nullsc after p7_bg_SetLength(bg, L), as p7_bg_NullOne computes it. */
float ref_nullsc(void *vh, int L)
{
  REFPROF *h = vh;
  float nullsc;
  p7_bg_SetLength(h->bg, L);
  p7_bg_NullOne(h->bg, NULL, L, &nullsc);
  return nullsc;
}

/* The MSV-stage pass decision of p7_Pipeline (p7_pipeline.c:667-677) for a
 * given usc: returns 1 when the target would pass the F1 filter. */
int ref_pass(void *vh, int L, float usc, double F1)
{
  REFPROF *h = vh;
  float nullsc, seq_score;
  double P;
  p7_bg_SetLength(h->bg, L);
  p7_bg_NullOne(h->bg, NULL, L, &nullsc);
  seq_score = (usc - nullsc) / eslCONST_LOG2;
  P = esl_gumbel_surv(seq_score, h->om->evparam[p7_MMU], h->om->evparam[p7_MLAMBDA]);
  return (P > F1) ? 0 : 1;
}

/* ---- batch MSV on N threads (CPU baseline) ---- */

typedef struct {
  REFPROF *h;
  const uint8_t *res;      /* concatenated residues, 1 byte each, codes 0..Kp-1 */
  const int64_t *off;      /* start of each target in res */
  const int32_t *len;
  float *usc;
  int32_t *status;
  int n;
  int next;                /* shared work counter */
  int vit;                 /* 0: p7_MSVFilter, 1: p7_ViterbiFilter */
  pthread_mutex_t lock;
} BATCH;

static void *batch_worker(void *arg)
{
  BATCH *b = arg;
  P7_OPROFILE *om = p7_oprofile_Clone(b->h->om);
  P7_OMX *ox = p7_omx_Create(om->M, 0, 0);
  int cap = 0;
  uint8_t *dsq = NULL;
  for (;;) {
    int i0, i1, i;
    pthread_mutex_lock(&b->lock);
    i0 = b->next; b->next += 256; if (b->next > b->n) b->next = b->n;
    i1 = b->next;
    pthread_mutex_unlock(&b->lock);
    if (i0 >= i1) break;
    for (i = i0; i < i1; i++) {
      int L = b->len[i];
      if (L + 2 > cap) { cap = L + 2; dsq = realloc(dsq, cap); }
      dsq[0] = dsq[L + 1] = eslDSQ_SENTINEL;
      memcpy(dsq + 1, b->res + b->off[i], L);
      p7_oprofile_ReconfigLength(om, L);
      b->status[i] = b->vit ? p7_ViterbiFilter((const ESL_DSQ *) dsq, L, om, ox, &b->usc[i])
                            : p7_MSVFilter((const ESL_DSQ *) dsq, L, om, ox, &b->usc[i]);
    }
  }
  free(dsq);
  p7_omx_Destroy(ox);
  p7_oprofile_Destroy(om);
  return NULL;
}

static int run_batch(void *vh, const uint8_t *res, const int64_t *off, const int32_t *len, int n,
                     float *usc, int32_t *status, int nthreads, int vit)
{
  BATCH b;
  pthread_t *tid;
  int t;
  if (nthreads < 1) nthreads = 1;
  b.h = vh; b.res = res; b.off = off; b.len = len; b.usc = usc; b.status = status;
  b.n = n; b.next = 0; b.vit = vit;
  pthread_mutex_init(&b.lock, NULL);
  tid = malloc(sizeof(pthread_t) * nthreads);
  for (t = 0; t < nthreads; t++) pthread_create(&tid[t], NULL, batch_worker, &b);
  for (t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
  free(tid);
  pthread_mutex_destroy(&b.lock);
  return 0;
}

int ref_msv_batch(void *vh, const uint8_t *res, const int64_t *off, const int32_t *len, int n,
                  float *usc, int32_t *status, int nthreads)
{
  return run_batch(vh, res, off, len, n, usc, status, nthreads, 0);
}

int ref_vit_batch(void *vh, const uint8_t *res, const int64_t *off, const int32_t *len, int n,
                  float *sc, int32_t *status, int nthreads)
{
  return run_batch(vh, res, off, len, n, sc, status, nthreads, 1);
}

/* ---- the ViterbiFilter stage ---- */

/* p7_ViterbiFilter on dsq[0..L+1] (digital, sentinels at 0 and L+1). */
int ref_vit(void *vh, const uint8_t *dsq, int L, float *sc)
{
  REFPROF *h = vh;
  p7_oprofile_ReconfigLength(h->om, L);
  return p7_ViterbiFilter((const ESL_DSQ *) dsq, L, h->om, h->ox, sc);
}

/* Q = p7O_NQW(M) of the striped Viterbi vectors. */
int ref_vit_Q(void *vh) { return p7O_NQW(((REFPROF *) vh)->om->M); }

/* The ViterbiFilter fields, de-striped over all 8*Q positions (position
 * p = q + z*Q holds node k = p+1; positions >= M are the padding):
 *   rwv[x * 8Q + p], twv[t * 8Q + p] with t = p7O_BM..p7O_DD (vitfilter.c's
 *   vector order: the 7 per-q transitions, then the DD run);
 *   ints = {xw[E][LOOP], xw[E][MOVE], xw[N][MOVE], xw[N][LOOP], xw[C][MOVE],
 *           xw[C][LOOP], xw[J][MOVE], xw[J][LOOP], base_w, ddbound_w}
 *          (the N/C/J moves for the length the profile was last set to);
 *   flts = {scale_w, nj, evparam[VMU], evparam[VLAMBDA]}. */
void ref_vit_fields(void *vh, int16_t *rwv, int16_t *twv, int *ints, float *flts)
{
  P7_OPROFILE *om = ((REFPROF *) vh)->om;
  const int Q = p7O_NQW(om->M), Kp = om->abc->Kp, W = 8 * Q;
  int x, q, z, t;
  union { __m128i v; int16_t i[8]; } u;
  for (x = 0; x < Kp; x++)
    for (q = 0; q < Q; q++) {
      u.v = om->rwv[x][q];
      for (z = 0; z < 8; z++) rwv[x * W + q + z * Q] = u.i[z];
    }
  for (q = 0; q < Q; q++)
    for (t = 0; t < 7; t++) {
      u.v = om->twv[q * 7 + t];
      for (z = 0; z < 8; z++) twv[t * W + q + z * Q] = u.i[z];
    }
  for (q = 0; q < Q; q++) {
    u.v = om->twv[7 * Q + q];
    for (z = 0; z < 8; z++) twv[7 * W + q + z * Q] = u.i[z];
  }
  ints[0] = om->xw[p7O_E][p7O_LOOP]; ints[1] = om->xw[p7O_E][p7O_MOVE];
  ints[2] = om->xw[p7O_N][p7O_MOVE]; ints[3] = om->xw[p7O_N][p7O_LOOP];
  ints[4] = om->xw[p7O_C][p7O_MOVE]; ints[5] = om->xw[p7O_C][p7O_LOOP];
  ints[6] = om->xw[p7O_J][p7O_MOVE]; ints[7] = om->xw[p7O_J][p7O_LOOP];
  ints[8] = om->base_w; ints[9] = om->ddbound_w;
  flts[0] = om->scale_w; flts[1] = om->nj;
  flts[2] = om->evparam[p7_VMU]; flts[3] = om->evparam[p7_VLAMBDA];
}

/* xw[N][MOVE] after p7_oprofile_ReconfigLength(om, L). */
int ref_vit_move(void *vh, int L)
{
  P7_OPROFILE *om = ((REFPROF *) vh)->om;
  p7_oprofile_ReconfigLength(om, L);
  return om->xw[p7O_N][p7O_MOVE];
}

/* Generic-profile transition scores: out[k * 8 + s] = p7P_TSC(gm, k, s) for
 * k = 0..M, s = p7P_MM, IM, DM, BM, MD, MI, II, DD (hmmer.h order). */
void ref_profile_tsc(void *vh, float *out)
{
  P7_PROFILE *gm = ((REFPROF *) vh)->gm;
  int k, s;
  for (k = 0; k <= gm->M; k++)
    for (s = 0; s < p7P_NTRANS; s++) out[k * p7P_NTRANS + s] = p7P_TSC(gm, k, s);
}

/* ---- generators, for pinning the product's restatement ---- */

/* n iid targets of length L from esl_randomness_Create(seed) drawn with
 * esl_rsq_xfIID over the amino background; out is n * L residues. */
void ref_iid(uint32_t seed, int n, int L, uint8_t *out)
{
  ESL_ALPHABET *abc = esl_alphabet_Create(eslAMINO);
  P7_BG *bg = p7_bg_Create(abc);
  ESL_RANDOMNESS *r = esl_randomness_Create(seed);
  ESL_DSQ *dsq = malloc(L + 2);
  int i;
  for (i = 0; i < n; i++) {
    esl_rsq_xfIID(r, bg->f, abc->K, L, dsq);
    memcpy(out + (size_t) i * L, dsq + 1, L);
  }
  free(dsq);
  esl_randomness_Destroy(r);
  p7_bg_Destroy(bg);
  esl_alphabet_Destroy(abc);
}

/* The first n uniform deviates of an RNG: fast = 0 for
 * esl_randomness_Create (Mersenne), 1 for esl_randomness_CreateFast. */
void ref_uniform(uint32_t seed, int fast, int n, double *out)
{
  ESL_RANDOMNESS *r = fast ? esl_randomness_CreateFast(seed) : esl_randomness_Create(seed);
  int i;
  for (i = 0; i < n; i++) out[i] = esl_random(r);
  esl_randomness_Destroy(r);
}

/* n core emissions from the profile's HMM using esl_randomness_Create(seed);
 * concatenated into out (capacity cap), lengths into lens.  Returns total
 * residues written, or -1 if cap is too small. */
int64_t ref_core_emit(void *vh, uint32_t seed, int n, uint8_t *out, int64_t cap, int32_t *lens)
{
  REFPROF *h = vh;
  ESL_RANDOMNESS *r = esl_randomness_Create(seed);
  ESL_SQ *sq = esl_sq_CreateDigital(h->abc);
  int64_t tot = 0;
  int i;
  for (i = 0; i < n; i++) {
    p7_CoreEmit(r, h->hmm, sq, NULL);
    if (tot + sq->n > cap) { tot = -1; break; }
    memcpy(out + tot, sq->dsq + 1, sq->n);
    lens[i] = (int32_t) sq->n;
    tot += sq->n;
  }
  esl_sq_Destroy(sq);
  esl_randomness_Destroy(r);
  return tot;
}

/* ---- real models from HMMER3 files (p7_hmmfile_OpenE / p7_hmmfile_Read,
 * src/p7_hmmfile.c), configured as hmmsearch does (p7_ProfileConfig L=100,
 * p7_LOCAL; cudaHmmsearch.cu:499-502).  The file's STATS lines provide the
 * calibrated MSV mu and lambda. ---- */
void *ref_profile_from_file(const char *path, int which)
{
  P7_HMMFILE *hfp = NULL;
  REFPROF *h = calloc(1, sizeof(REFPROF));
  int i;
  if (p7_hmmfile_OpenE(path, NULL, &hfp, NULL) != eslOK) { free(h); return NULL; }
  for (i = 0; i <= which; i++) {
    if (h->hmm) { p7_hmm_Destroy(h->hmm); h->hmm = NULL; }
    if (p7_hmmfile_Read(hfp, &h->abc, &h->hmm) != eslOK) { p7_hmmfile_Close(hfp); refprof_free(h); return NULL; }
  }
  p7_hmmfile_Close(hfp);
  if (refprof_configure(h, 100, 0) != 0) { refprof_free(h); return NULL; }
  return h;
}

int ref_profile_Kp(void *vh) { return ((REFPROF *) vh)->abc->Kp; }

/* Digitize text with the profile's alphabet (esl_abc_Digitize); out gets
 * codes for positions 1..n (no sentinels).  Returns n, or -1 on bad input. */
int ref_digitize(void *vh, const char *text, uint8_t *out)
{
  REFPROF *h = vh;
  int n = (int) strlen(text);
  ESL_DSQ *dsq = malloc(n + 2);
  if (esl_abc_Digitize(h->abc, text, dsq) != eslOK) { free(dsq); return -1; }
  memcpy(out, dsq + 1, n);
  free(dsq);
  return n;
}
