/* p7_MSVFilter() and p7_ViterbiFilter() for HMMER 3.1b1's impl_sse backend,
 * computed on an MI355X through libmsvgpu.
 *
 * Replaces the functions of src/impl_sse/msvfilter.c:78-213 and
 * src/impl_sse/vitfilter.c:84-250 (declared at src/impl_sse/impl_sse.h:327
 * and :343) with the same signatures, the same status codes (eslOK; eslERANGE
 * with *ret_sc = +inf; eslEINVAL through ESL_EXCEPTION) and bit-identical
 * scores.  Every caller (p7_Pipeline's first two filters,
 * p7_pipeline.c:671 and :701, p7_MSVMu, p7_ViterbiMu, the drivers) takes
 * them unchanged.
 *
 * Build: compile this file against the reference headers and the include/
 * directory of this package, link it ahead of the impl_sse objects with the
 * stock objects' own symbols renamed
 *     objcopy --redefine-sym p7_MSVFilter=p7_MSVFilter_sse msvfilter.o
 *     objcopy --redefine-sym p7_ViterbiFilter=p7_ViterbiFilter_sse vitfilter.o
 * (the rest of those files -- the nhmmer long-target filters among it --
 * stays as it is), and link -lmsvgpu.
 *
 * Each host thread keeps its own context and its own uploaded profile, as
 * each hmmsearch worker keeps its own P7_OPROFILE clone.  The uploaded
 * profile is recognised by the contents of its MSV and ViterbiFilter fields,
 * not by the pointer: between queries a freed P7_OPROFILE and its successor
 * can share an address.  The length-dependent fields (tjb_b, xw[N|C|J][MOVE])
 * are derived from L by the library with the expressions
 * p7_oprofile_ReconfigLength() uses (p7_oprofile.c:897-962); an om whose
 * fields are not those of L (a length model elided or set for another L, as
 * hmmsim --x-no-lengthmodel does) is refused with eslEINVAL rather than
 * scored differently from the SSE filters.  This path
 * makes one host-device round trip per target; msvg_search and
 * msvg_vit_search are the batch paths.
 *
 * p7_mi355x_PipelineBlock is the batch path inside HMMER: the block step of
 * cudaHmmsearch (src/GPUFunction.cu:360-490: the GPU scores a block of
 * sequences, then p7_PipelineShort, src/p7_pipeline.c:887-1000, runs the
 * rest of p7_Pipeline on the host for each with its precomputed MSV score),
 * with the ViterbiFilter of the block's MSV survivors scored on the GPU as
 * well.  p7_PipelineShort calls p7_ViterbiFilter itself; inside the block
 * step that call takes the precomputed score of its sequence.
 *
 * p7_mi355x_ThreadLoop is hmmsearch's thread_loop with pipeline_thread
 * (src/hmmsearch.c:1317-1425; cudaHmmsearch.cu:670-800 drives its GPU
 * worker the same way): the calling thread reads the database in blocks
 * while worker threads each take a block and run the block step on it, so
 * one block's GPU filters, another's host tail (p7_PipelineShort on its
 * survivors) and the reading of the next overlap.  A worker thread borrows
 * a device context from a pool kept for the process (the contexts live on
 * between queries), so a worker's set-up costs a profile upload, not a
 * device initialisation.
 */
#include "p7_config.h"

#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "easel.h"
#include "hmmer.h"

#include "msv_gpu.h"
#include "msv_prep.h"
#include "filters_mi355x.h"

struct mi355x_state {
  msvg_ctx *ctx;
  int32_t   pid;      /* uploaded profile, or -1                              */
  uint8_t  *key;      /* the uploaded profile's fields, byte for byte         */
  size_t    keylen, keycap;
  uint8_t  *tmp;      /* the same for the profile of the current call         */
  size_t    tmpcap;
  /* inside p7_mi355x_PipelineBlock: the ViterbiFilter scores of the block's
   * MSV survivors, in block order                                            */
  int             in_block;
  const ESL_DSQ **vdsq;
  int64_t        *vlen;
  float          *vsc;
  uint8_t        *vst;
  int64_t         nv, vnext;
  struct mi355x_state *next;   /* in the pool of idle worker states          */
};

/* The calling thread's state: its own (created on first use, kept for the
 * thread's life) or, inside a p7_mi355x_ThreadLoop worker, one borrowed
 * from the pool. */
static __thread struct mi355x_state *g_cur = NULL;
static pthread_mutex_t      g_pool_mu = PTHREAD_MUTEX_INITIALIZER;
static struct mi355x_state *g_pool = NULL;

static struct mi355x_state *new_state(void)
{
  struct mi355x_state *s = calloc(1, sizeof *s);
  if (s) s->pid = -1;
  return s;
}

static struct mi355x_state *cur_state(void)
{
  if (g_cur == NULL) g_cur = new_state();
  return g_cur;
}

static struct mi355x_state *pool_take(void)
{
  struct mi355x_state *s;
  pthread_mutex_lock(&g_pool_mu);
  if ((s = g_pool) != NULL) g_pool = s->next;
  pthread_mutex_unlock(&g_pool_mu);
  return s ? s : new_state();
}

static void pool_put(struct mi355x_state *s)
{
  pthread_mutex_lock(&g_pool_mu);
  s->next = g_pool;
  g_pool  = s;
  pthread_mutex_unlock(&g_pool_mu);
}

static uint8_t *put(uint8_t *p, const void *src, size_t n) { memcpy(p, src, n); return p + n; }

/* The bytes that decide both filters' results: M, Kp, mode, nj, the MSV
 * scalars and rbv stripes, the ViterbiFilter scalars, rwv and twv stripes,
 * and the calibrated MSV and Viterbi E-value parameters.  Row strides are
 * the allocation's, so rows are taken one at a time. */
static int make_key(const P7_OPROFILE *om, uint8_t **buf, size_t *ret_len, size_t *cap)
{
  const int Qb = p7O_NQB(om->M), Qw = p7O_NQW(om->M), Kp = om->abc->Kp;
  const size_t head = 3 * sizeof(int) + 5 * sizeof(float) + 4 + 4 * sizeof(int16_t);
  const size_t len = head + (size_t) Kp * Qb * 16 + (size_t) Kp * Qw * 16 + (size_t) 8 * Qw * 16;
  uint8_t *p;
  int x;

  if (len > *cap) {
    if ((p = realloc(*buf, len)) == NULL) return eslEMEM;
    *buf = p;
    *cap = len;
  }
  p = *buf;
  p = put(p, &om->M, sizeof(int));
  p = put(p, &Kp, sizeof(int));
  p = put(p, &om->mode, sizeof(int));
  p = put(p, &om->nj, sizeof(float));
  *p++ = om->tbm_b; *p++ = om->tec_b; *p++ = om->base_b; *p++ = om->bias_b;
  p = put(p, &om->scale_b, sizeof(float));
  p = put(p, &om->xw[p7O_E][p7O_LOOP], sizeof(int16_t));
  p = put(p, &om->xw[p7O_E][p7O_MOVE], sizeof(int16_t));
  p = put(p, &om->base_w, sizeof(int16_t));
  p = put(p, &om->ddbound_w, sizeof(int16_t));
  p = put(p, &om->scale_w, sizeof(float));
  p = put(p, &om->evparam[p7_MMU], sizeof(float));
  p = put(p, &om->evparam[p7_MLAMBDA], sizeof(float));
  for (x = 0; x < Kp; x++) p = put(p, om->rbv[x], (size_t) Qb * 16);
  for (x = 0; x < Kp; x++) p = put(p, om->rwv[x], (size_t) Qw * 16);
  p = put(p, om->twv, (size_t) 8 * Qw * 16);
  *ret_len = len;
  return eslOK;
}

/* De-stripe the MSV and ViterbiFilter vectors into the model-order tables of
 * the ABI: rbv byte z of vector q is node k = q + z*Q + 1 (p7_oprofile.c:
 * 625-630); rwv/twv word z of vector q likewise with Q = p7O_NQW(M), twv
 * holding 7 vectors per q followed by the DD run (p7_oprofile.c:680-715). */
static int upload(struct mi355x_state *s, const P7_OPROFILE *om)
{
  const int M = om->M, Qb = p7O_NQB(M), Qw = p7O_NQW(M), Kp = om->abc->Kp;
  msvg_oprofile d;
  msvg_vprofile v;
  uint8_t *rbv = malloc((size_t) Kp * M);
  int16_t *rwv = malloc(sizeof(int16_t) * (size_t) Kp * M);
  int16_t *twv = malloc(sizeof(int16_t) * (size_t) 8 * M);
  int x, q, z, t, status;

  if (!rbv || !rwv || !twv) { free(rbv); free(rwv); free(twv); return eslEMEM; }
  for (x = 0; x < Kp; x++) {
    const uint8_t *brow = (const uint8_t *) om->rbv[x];
    const int16_t *wrow = (const int16_t *) om->rwv[x];
    for (q = 0; q < Qb; q++)
      for (z = 0; z < 16; z++) {
        const int k = q + z * Qb + 1;
        if (k <= M) rbv[(size_t) x * M + k - 1] = brow[q * 16 + z];
      }
    for (q = 0; q < Qw; q++)
      for (z = 0; z < 8; z++) {
        const int k = q + z * Qw + 1;
        if (k <= M) rwv[(size_t) x * M + k - 1] = wrow[q * 8 + z];
      }
  }
  for (q = 0; q < Qw; q++)
    for (z = 0; z < 8; z++) {
      const int k = q + z * Qw + 1;
      if (k > M) continue;
      for (t = 0; t < 7; t++) twv[(size_t) t * M + k - 1] = ((const int16_t *) (om->twv + q * 7 + t))[z];
      twv[(size_t) 7 * M + k - 1] = ((const int16_t *) (om->twv + 7 * Qw + q))[z];
    }

  d.M = M;             d.Kp = Kp;            d.rbv = rbv;
  d.tbm_b = om->tbm_b; d.tec_b = om->tec_b;  d.base_b = om->base_b; d.bias_b = om->bias_b;
  d.scale_b = om->scale_b;
  d.mmu = om->evparam[p7_MMU];
  d.mlambda = om->evparam[p7_MLAMBDA];
  v.M = M; v.Kp = Kp; v.rwv = rwv; v.twv = twv;
  v.e_loop = om->xw[p7O_E][p7O_LOOP]; v.e_move = om->xw[p7O_E][p7O_MOVE];
  v.base_w = om->base_w; v.ddbound_w = om->ddbound_w;
  v.scale_w = om->scale_w; v.nj = om->nj;
  v.vmu = om->evparam[p7_VMU]; v.vlambda = om->evparam[p7_VLAMBDA];

  if (s->pid >= 0) msvg_profile_remove(s->ctx, s->pid);
  s->pid = -1;
  status = msvg_profile_add(s->ctx, &d, &s->pid);
  if (status == MSVG_OK) status = msvg_profile_set_viterbi(s->ctx, s->pid, &v);
  if (status != MSVG_OK && s->pid >= 0) { msvg_profile_remove(s->ctx, s->pid); s->pid = -1; }
  free(rbv); free(rwv); free(twv);
  return status;
}

/* Context and uploaded profile for om; eslOK or an exception. */
static int prepare(struct mi355x_state *s, const P7_OPROFILE *om)
{
  size_t keylen = 0;
  int status;
  if (s->ctx == NULL && msvg_init(0, &s->ctx) != MSVG_OK)
    ESL_EXCEPTION(eslEINVAL, "MI355X filters: no MI355X device available");
  if ((status = make_key(om, &s->tmp, &keylen, &s->tmpcap)) != eslOK) ESL_EXCEPTION(status, "allocation failed");
  if (s->pid < 0 || keylen != s->keylen || memcmp(s->tmp, s->key, keylen) != 0) {
    uint8_t *k = s->key;
    size_t cap = s->keycap;
    if ((status = upload(s, om)) != MSVG_OK) ESL_EXCEPTION(status, "MI355X filters: %s", msvg_last_error(s->ctx));
    s->key = s->tmp;  s->keycap = s->tmpcap;  s->keylen = keylen;
    s->tmp = k;       s->tmpcap = cap;
  }
  return eslOK;
}

int p7_MSVFilter(const ESL_DSQ *dsq, int L, const P7_OPROFILE *om, P7_OMX *ox, float *ret_sc)
{
  struct mi355x_state *s = cur_state();
  int status;

  if (p7O_NQB(om->M) > ox->allocQ16) ESL_EXCEPTION(eslEINVAL, "DP matrix allocated too small");  /* msvfilter.c:103 */
  ox->M = om->M;                                                                             /* :104 */
  if (om->tjb_b != (uint8_t) msvh_tjb(L))                                                    /* :120 */
    ESL_EXCEPTION(eslEINVAL, "p7_MSVFilter: om's length model is not that of L = %d (p7_oprofile_ReconfigLength)", L);
  if (s == NULL) ESL_EXCEPTION(eslEMEM, "allocation failed");
  if ((status = prepare(s, om)) != eslOK) return status;
  status = msvg_MSVFilter(s->ctx, s->pid, dsq, L, ret_sc);
  if (status != MSVG_OK && status != MSVG_ERANGE)
    ESL_EXCEPTION(status, "p7_MSVFilter: %s", msvg_last_error(s->ctx));
  return status;
}

int p7_ViterbiFilter(const ESL_DSQ *dsq, int L, const P7_OPROFILE *om, P7_OMX *ox, float *ret_sc)
{
  struct mi355x_state *s = cur_state();
  int status;

  if (p7O_NQW(om->M) > ox->allocQ8) ESL_EXCEPTION(eslEINVAL, "DP matrix allocated too small");  /* vitfilter.c:104 */
  if (om->mode != p7_LOCAL && om->mode != p7_UNILOCAL)                                          /* :105 */
    ESL_EXCEPTION(eslEINVAL, "Fast filter only works for local alignment");
  ox->M = om->M;                                                                                /* :106 */
  {                                          /* the N/C/J costs of vitfilter.c:116-117, :179-181, :241 */
    const int move = msvh_vit_move(om->scale_w, om->nj, L);
    if (om->xw[p7O_N][p7O_MOVE] != move || om->xw[p7O_C][p7O_MOVE] != move || om->xw[p7O_J][p7O_MOVE] != move ||
        om->xw[p7O_N][p7O_LOOP] != 0 || om->xw[p7O_C][p7O_LOOP] != 0 || om->xw[p7O_J][p7O_LOOP] != 0)
      ESL_EXCEPTION(eslEINVAL, "p7_ViterbiFilter: om's length model is not that of L = %d (p7_oprofile_ReconfigLength)", L);
  }
  if (s == NULL) ESL_EXCEPTION(eslEMEM, "allocation failed");
  if (s->in_block) {               /* a survivor of the current block: its precomputed score */
    int64_t k = s->vnext;
    while (k < s->nv && s->vdsq[k] != dsq) k++;
    if (k == s->nv) for (k = 0; k < s->nv && s->vdsq[k] != dsq; k++) ;
    if (k < s->nv && s->vlen[k] == L) {
      s->vnext = k + 1;
      *ret_sc = s->vsc[k];
      return s->vst[k] ? eslERANGE : eslOK;
    }
  }
  if ((status = prepare(s, om)) != eslOK) return status;
  status = msvg_ViterbiFilter(s->ctx, s->pid, dsq, L, ret_sc);
  if (status != MSVG_OK && status != MSVG_ERANGE)
    ESL_EXCEPTION(status, "p7_ViterbiFilter: %s", msvg_last_error(s->ctx));
  return status;
}

int p7_mi355x_PipelineBlock(P7_PIPELINE *pli, P7_OPROFILE *om, P7_BG *bg, ESL_SQ **sq, int n, P7_TOPHITS *th)
{
  struct mi355x_state *s = cur_state();
  uint8_t  *res = NULL, *st = NULL, *ps = NULL, *vst = NULL;
  int64_t  *off = NULL, *tgt = NULL, *vlen = NULL;
  int32_t  *len = NULL, did = -1;
  float    *usc = NULL, *vsc = NULL;
  const ESL_DSQ **vdsq = NULL;
  int64_t   total = 0, nv = 0;
  int       i, gpu = MSVG_OK, status = eslOK;

  if (n <= 0) return eslOK;
  if (s == NULL) ESL_EXCEPTION(eslEMEM, "allocation failed");
  if ((status = prepare(s, om)) != eslOK) return status;
  for (i = 0; i < n; i++) total += sq[i]->n;
  res = malloc(total > 0 ? (size_t) total : 1);
  off = malloc(sizeof(int64_t) * n);   len = malloc(sizeof(int32_t) * n);
  usc = malloc(sizeof(float) * n);     st = malloc(n);   ps = malloc(n);
  tgt = malloc(sizeof(int64_t) * n);   vsc = malloc(sizeof(float) * n);   vst = malloc(n);
  vdsq = malloc(sizeof(ESL_DSQ *) * n);  vlen = malloc(sizeof(int64_t) * n);
  if (!res || !off || !len || !usc || !st || !ps || !tgt || !vsc || !vst || !vdsq || !vlen) { gpu = MSVG_EMEM; goto DONE; }
  total = 0;
  for (i = 0; i < n; i++) {        /* digital residues without the sentinels */
    off[i] = total;
    len[i] = (int32_t) sq[i]->n;
    if (sq[i]->n > 0) memcpy(res + total, sq[i]->dsq + 1, (size_t) sq[i]->n);
    total += sq[i]->n;
  }
  if ((gpu = msvg_db_add(s->ctx, res, off, len, n, &did)) != MSVG_OK) goto DONE;
  /* eager: the block's database is searched once, so no graph is captured */
  if ((gpu = msvg_search(s->ctx, &s->pid, 1, did, pli->F1, MSVG_MODE_BATCHED | MSVG_FLAG_EAGER, usc, st, ps)) != MSVG_OK)
    goto DONE;
  for (i = 0; i < n; i++)          /* the MSV survivors: p7_PipelineShort's first test, p7_pipeline.c:917-919 */
    if (ps[i] && sq[i]->n > 0) tgt[nv++] = i;
  if (nv > 0 && (gpu = msvg_vit_search(s->ctx, s->pid, did, tgt, nv, vsc, vst)) != MSVG_OK) goto DONE;
  for (i = 0; i < nv; i++) { vdsq[i] = sq[tgt[i]]->dsq; vlen[i] = sq[tgt[i]]->n; }
  s->vdsq = vdsq; s->vlen = vlen; s->vsc = vsc; s->vst = vst; s->nv = nv; s->vnext = 0;
  s->in_block = 1;
  for (i = 0; i < n && status == eslOK; i++) {     /* GPUFunction.cu:474-488 */
    p7_pli_NewSeq(pli, sq[i]);
    /* p7_PipelineShort returns at its MSV test (p7_pipeline.c:917-919) for
     * the sequences whose pass bit is 0, with nothing changed: the same
     * F1 decision on the same usc */
    if (!ps[i]) continue;
    p7_bg_SetLength(bg, sq[i]->n);
    p7_oprofile_ReconfigLength(om, sq[i]->n);
    status = p7_PipelineShort(pli, om, bg, sq[i], th, usc[i]);
    p7_pipeline_Reuse(pli);
  }
  s->in_block = 0;
  s->nv = 0;
 DONE:
  if (did >= 0) msvg_db_remove(s->ctx, did);
  free(res); free(off); free(len); free(usc); free(st); free(ps); free(tgt); free(vsc); free(vst);
  free(vdsq); free(vlen);
  if (gpu != MSVG_OK) ESL_EXCEPTION(gpu, "p7_mi355x_PipelineBlock: %s", msvg_last_error(s->ctx));
  return status;
}

int p7_mi355x_PipelineSqBlock(P7_PIPELINE *pli, P7_OPROFILE *om, P7_BG *bg, ESL_SQ_BLOCK *block, P7_TOPHITS *th)
{
  ESL_SQ **sq;
  int      i, status;

  if (block->count <= 0) return eslOK;
  if ((sq = malloc(sizeof(ESL_SQ *) * block->count)) == NULL) ESL_EXCEPTION(eslEMEM, "allocation failed");
  for (i = 0; i < block->count; i++) sq[i] = block->list + i;
  status = p7_mi355x_PipelineBlock(pli, om, bg, sq, block->count, th);
  for (i = 0; i < block->count; i++) esl_sq_Reuse(block->list + i);   /* hmmsearch.c:1410 */
  free(sq);
  return status;
}

/* The work queue between the reader and the workers: blocks cycle from
 * idle (the reader fills one) to full (a worker takes one) and back. */
struct tl_queue {
  pthread_mutex_t mu;
  pthread_cond_t  filled;      /* a full block, or the end of the input    */
  pthread_cond_t  emptied;     /* an idle block, or a failure              */
  ESL_SQ_BLOCK  **full, **idle;
  int             nfull, nidle, head, cap;
  int             done;        /* the reader has queued its last block     */
  int             status;      /* eslOK, or the first worker failure       */
};

struct tl_worker {
  struct tl_queue  *q;
  P7_MI355X_WORKER *w;
  pthread_t         tid;
};

static void *tl_work(void *arg)
{
  struct tl_worker    *tw = arg;
  struct tl_queue     *q = tw->q;
  struct mi355x_state *s = pool_take();
  ESL_SQ_BLOCK        *b;
  int                  st, i;

  g_cur = s;
  for (;;) {
    pthread_mutex_lock(&q->mu);
    while (q->nfull == 0 && !q->done) pthread_cond_wait(&q->filled, &q->mu);
    if (q->nfull == 0) { pthread_mutex_unlock(&q->mu); break; }
    b = q->full[q->head];
    q->head = (q->head + 1) % q->cap;
    q->nfull--;
    st = q->status;
    pthread_mutex_unlock(&q->mu);

    if (st == eslOK && s == NULL) st = eslEMEM;
    if (st == eslOK) st = p7_mi355x_PipelineSqBlock(tw->w->pli, tw->w->om, tw->w->bg, b, tw->w->th);
    else for (i = 0; i < b->count; i++) esl_sq_Reuse(b->list + i);      /* after a failure: drained */

    pthread_mutex_lock(&q->mu);
    q->idle[q->nidle++] = b;
    if (st != eslOK && q->status == eslOK) q->status = st;
    pthread_cond_signal(&q->emptied);
    pthread_mutex_unlock(&q->mu);
  }
  g_cur = NULL;
  if (s) pool_put(s);
  return NULL;
}

int p7_mi355x_ThreadLoop(P7_MI355X_WORKER *info, int nworkers, ESL_SQFILE *dbfp, int block_seqs, int n_targetseqs)
{
  struct tl_queue   q;
  struct tl_worker *tw = NULL;
  int               nblk, started = 0, i, sstatus = eslOK;

  if (nworkers < 1 || nworkers > 256 || block_seqs < 1) ESL_EXCEPTION(eslEINVAL, "p7_mi355x_ThreadLoop: bad arguments");
  memset(&q, 0, sizeof q);
  q.status = eslOK;
  nblk = q.cap = 2 * nworkers + 1;        /* every worker busy, one more full each, and the one being read */
  pthread_mutex_init(&q.mu, NULL);
  pthread_cond_init(&q.filled, NULL);
  pthread_cond_init(&q.emptied, NULL);
  q.full = calloc(nblk, sizeof(ESL_SQ_BLOCK *));
  q.idle = calloc(nblk, sizeof(ESL_SQ_BLOCK *));
  tw     = calloc(nworkers, sizeof *tw);
  if (!q.full || !q.idle || !tw) { sstatus = eslEMEM; goto DONE; }
  for (i = 0; i < nblk; i++) {
    if ((q.idle[i] = esl_sq_CreateDigitalBlock(block_seqs, info[0].om->abc)) == NULL) { sstatus = eslEMEM; goto DONE; }
    q.nidle++;
  }
  for (i = 0; i < nworkers; i++, started++) {
    tw[i].q = &q;
    tw[i].w = &info[i];
    if (pthread_create(&tw[i].tid, NULL, tl_work, &tw[i]) != 0) { sstatus = eslESYS; break; }
  }

  /* the reader: thread_loop, hmmsearch.c:1317-1369 */
  while (sstatus == eslOK) {
    ESL_SQ_BLOCK *b;
    pthread_mutex_lock(&q.mu);
    while (q.nidle == 0 && q.status == eslOK) pthread_cond_wait(&q.emptied, &q.mu);
    if (q.status != eslOK) { pthread_mutex_unlock(&q.mu); break; }
    b = q.idle[--q.nidle];
    pthread_mutex_unlock(&q.mu);

    if (n_targetseqs == 0) { b->count = 0; sstatus = eslEOF; }
    else {
      sstatus = esl_sqio_ReadBlock(dbfp, b, -1, n_targetseqs, FALSE);
      if (sstatus == eslOK && n_targetseqs > 0) n_targetseqs -= b->count;
    }

    pthread_mutex_lock(&q.mu);
    if (sstatus == eslOK) {
      q.full[(q.head + q.nfull) % q.cap] = b;
      q.nfull++;
      pthread_cond_signal(&q.filled);
    } else q.idle[q.nidle++] = b;
    pthread_mutex_unlock(&q.mu);
  }

 DONE:
  pthread_mutex_lock(&q.mu);
  q.done = 1;
  pthread_cond_broadcast(&q.filled);
  pthread_mutex_unlock(&q.mu);
  for (i = 0; i < started; i++) pthread_join(tw[i].tid, NULL);
  if (q.idle) for (i = 0; i < q.nidle; i++) esl_sq_DestroyBlock(q.idle[i]);
  free(q.full); free(q.idle); free(tw);
  pthread_cond_destroy(&q.filled);
  pthread_cond_destroy(&q.emptied);
  pthread_mutex_destroy(&q.mu);
  if (q.status != eslOK) return q.status;
  return sstatus;
}
