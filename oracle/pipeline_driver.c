/* pipeline_driver -- test infrastructure: a serial hmmsearch over the
 * reference's own library, linked twice (oracle/Makefile `dropin`):
 *
 *   oracle/_ref/pipeline_ref      with the reference's SSE p7_MSVFilter and
 *                                 p7_ViterbiFilter
 *   oracle/_ref/pipeline_mi355x   with cudahmmsearch_amd/hmmer/filters_mi355x.c
 *                                 (both filters on the GPU) in their place
 *
 * Identical output files from the two are the drop-in check: every
 * p7_MSVFilter and p7_ViterbiFilter call made by p7_Pipeline goes through the
 * replaced symbols, so the filter pass counts, the hits that survive the
 * later stages and every per-target filter score below must agree bit for
 * bit.
 *
 * The loop is hmmsearch's serial_loop (src/hmmsearch.c:1285-1313) with the
 * per-query set-up of src/hmmsearch.c:474-510 and the report calls of
 * src/hmmsearch.c:549-559; the stopwatch is left out so the output holds no
 * timing.  Own code; only reference functions are called.
 *
 *   oracle/_ref/pipeline_block    the same library as pipeline_mi355x, the
 *                                 database searched by the block step,
 *                                 p7_mi355x_PipelineBlock (both filters
 *                                 batched on the GPU, then p7_PipelineShort):
 *                                 with workers = 0 as one block in the main
 *                                 thread; with workers > 0 by
 *                                 p7_mi355x_ThreadLoop in blocks of
 *                                 block_seqs sequences over that many worker
 *                                 threads, their hit lists and pipelines
 *                                 merged as hmmsearch merges its workers'
 *                                 (src/hmmsearch.c:505-542)
 *
 *   pipeline_driver <hmmfile> <seqfile> <outfile> [F1|- [nobias|bias [nolines|lines [workers [block_seqs]]]]]
 *
 * In the SSE build, workers > 0 (with nolines) is hmmsearch --cpu workers:
 * a reader thread and that many threads running p7_Pipeline, merged.
 * badlength-msv / badlength-vit in place of nolines: before the first
 * target, call that filter with om configured for 3L + 100 (the GPU builds
 * refuse it through Easel's exception handler).
 *
 * After each query's report it lists, for every target, p7_MSVFilter's and
 * p7_ViterbiFilter's status and the float bits of their scores (not with
 * nolines).  The seconds spent in the search of each query, sequence reading
 * included, go to stderr.
 */
#include "p7_config.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <time.h>

#include "easel.h"
#include "esl_alphabet.h"
#include "esl_sq.h"
#include "esl_sqio.h"
#include "hmmer.h"
#ifdef MI355X_BLOCK
#include "filters_mi355x.h"
#else
#include <pthread.h>
#endif

static double now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double) t.tv_sec + 1e-9 * (double) t.tv_nsec;
}

static void die(const char *msg, const char *arg)
{
  fprintf(stderr, "pipeline_driver: %s %s\n", msg, arg ? arg : "");
  exit(1);
}

#ifndef MI355X_BLOCK
/* hmmsearch's threaded search for the SSE build (hmmsearch --cpu N,
 * src/hmmsearch.c:1317-1425): the main thread reads blocks of sequences,
 * worker threads run p7_Pipeline on every sequence of a block with their own
 * pipeline, profile clone, background and hit list (WORKER_INFO, :36-44);
 * the caller merges them (:538-542).  The reference's own thread_loop
 * needs esl_threads.c, which this image cannot build. */
typedef struct { P7_PIPELINE *pli; P7_OPROFILE *om; P7_BG *bg; P7_TOPHITS *th; pthread_t tid; void *q; } sse_worker;
typedef struct {
  pthread_mutex_t mu;
  pthread_cond_t  filled, emptied;
  ESL_SQ_BLOCK  **full, **idle;
  int             nfull, nidle, head, cap, done;
} sse_queue;

static void *sse_work(void *arg)
{
  sse_worker   *w = arg;
  sse_queue    *q = w->q;
  ESL_SQ_BLOCK *b;
  int           i;
  for (;;) {
    pthread_mutex_lock(&q->mu);
    while (q->nfull == 0 && !q->done) pthread_cond_wait(&q->filled, &q->mu);
    if (q->nfull == 0) { pthread_mutex_unlock(&q->mu); break; }
    b = q->full[q->head];
    q->head = (q->head + 1) % q->cap;
    q->nfull--;
    pthread_mutex_unlock(&q->mu);
    for (i = 0; i < b->count; i++) {          /* hmmsearch.c:1397-1410 */
      ESL_SQ *x = b->list + i;
      p7_pli_NewSeq(w->pli, x);
      p7_bg_SetLength(w->bg, x->n);
      p7_oprofile_ReconfigLength(w->om, x->n);
      p7_Pipeline(w->pli, w->om, w->bg, x, w->th);
      esl_sq_Reuse(x);
      p7_pipeline_Reuse(w->pli);
    }
    pthread_mutex_lock(&q->mu);
    q->idle[q->nidle++] = b;
    pthread_cond_signal(&q->emptied);
    pthread_mutex_unlock(&q->mu);
  }
  return NULL;
}

/* Returns the reader's last status: eslEOF after the whole file. */
static int sse_threaded(sse_worker *w, int nw, ESL_SQFILE *sqfp, const ESL_ALPHABET *abc, int block_seqs)
{
  sse_queue q;
  int       i, sstatus = eslOK;
  memset(&q, 0, sizeof q);
  q.cap = 2 * nw + 1;
  pthread_mutex_init(&q.mu, NULL);
  pthread_cond_init(&q.filled, NULL);
  pthread_cond_init(&q.emptied, NULL);
  q.full = calloc(q.cap, sizeof(ESL_SQ_BLOCK *));
  q.idle = calloc(q.cap, sizeof(ESL_SQ_BLOCK *));
  if (!q.full || !q.idle) die("out of memory", NULL);
  for (i = 0; i < q.cap; i++)
    if ((q.idle[q.nidle++] = esl_sq_CreateDigitalBlock(block_seqs, abc)) == NULL) die("out of memory", NULL);
  for (i = 0; i < nw; i++) {
    w[i].q = &q;
    if (pthread_create(&w[i].tid, NULL, sse_work, &w[i]) != 0) die("cannot start a worker thread", NULL);
  }
  while (sstatus == eslOK) {
    ESL_SQ_BLOCK *b;
    pthread_mutex_lock(&q.mu);
    while (q.nidle == 0) pthread_cond_wait(&q.emptied, &q.mu);
    b = q.idle[--q.nidle];
    pthread_mutex_unlock(&q.mu);
    sstatus = esl_sqio_ReadBlock(sqfp, b, -1, -1, FALSE);
    pthread_mutex_lock(&q.mu);
    if (sstatus == eslOK) {
      q.full[(q.head + q.nfull) % q.cap] = b;
      q.nfull++;
      pthread_cond_signal(&q.filled);
    } else q.idle[q.nidle++] = b;
    pthread_mutex_unlock(&q.mu);
  }
  pthread_mutex_lock(&q.mu);
  q.done = 1;
  pthread_cond_broadcast(&q.filled);
  pthread_mutex_unlock(&q.mu);
  for (i = 0; i < nw; i++) pthread_join(w[i].tid, NULL);
  for (i = 0; i < q.nidle; i++) esl_sq_DestroyBlock(q.idle[i]);
  free(q.full);
  free(q.idle);
  pthread_cond_destroy(&q.filled);
  pthread_cond_destroy(&q.emptied);
  pthread_mutex_destroy(&q.mu);
  return sstatus;
}
#endif

int main(int argc, char **argv)
{
  char         errbuf[eslERRBUFSIZE];
  P7_HMMFILE  *hfp = NULL;
  ESL_SQFILE  *sqfp = NULL;
  ESL_ALPHABET *abc = NULL;
  P7_HMM      *hmm = NULL;
  P7_BG       *bg = NULL;
  FILE        *ofp;
  double       F1 = -1.0;
  int          nobias = 0, nolines = 0, nquery = 0, hstatus, sstatus;
  int          workers = 0, block_seqs = 10000;
  int          badlength = 0;  /* test of the filters' length-model check: 1 MSV, 2 ViterbiFilter */

  if (argc < 4) die("usage: pipeline_driver <hmmfile> <seqfile> <outfile> "
                    "[F1|- [nobias|bias [nolines|lines [workers [block_seqs]]]]]", NULL);
  if (argc > 4 && strcmp(argv[4], "-") != 0) F1 = atof(argv[4]);
  if (argc > 5) nobias = (strcmp(argv[5], "nobias") == 0);
  if (argc > 6) nolines = (strcmp(argv[6], "nolines") == 0);
  if (argc > 6 && strncmp(argv[6], "badlength-", 10) == 0)
    badlength = (strcmp(argv[6] + 10, "msv") == 0) ? 1 : (strcmp(argv[6] + 10, "vit") == 0) ? 2 : 0;
  if (argc > 7) workers = atoi(argv[7]);
  if (argc > 8) block_seqs = atoi(argv[8]);
  if (workers < 0 || block_seqs < 1) die("bad workers / block_seqs", NULL);
  if (p7_hmmfile_OpenE(argv[1], NULL, &hfp, errbuf) != eslOK) die("cannot open HMM file", argv[1]);
  if ((ofp = fopen(argv[3], "w")) == NULL) die("cannot write", argv[3]);

  while ((hstatus = p7_hmmfile_Read(hfp, &abc, &hmm)) == eslOK) {
    P7_PROFILE  *gm = p7_profile_Create(hmm->M, abc);
    P7_OPROFILE *om = p7_oprofile_Create(hmm->M, abc);
    P7_OPROFILE *wom;
    P7_PIPELINE *pli;
    P7_TOPHITS  *th = p7_tophits_Create();
    P7_OMX      *ox;
    ESL_SQ      *sq;
    int64_t      nseq = 0;
    char        *lines = NULL;
    size_t       nlines = 0, cap = 0;

    if (nquery++ == 0) {
      bg = p7_bg_Create(abc);
      if (esl_sqfile_OpenDigital(abc, argv[2], eslSQFILE_UNKNOWN, NULL, &sqfp) != eslOK) die("cannot open", argv[2]);
    } else if (esl_sqfile_Position(sqfp, 0) != eslOK) die("cannot rewind", argv[2]);

    fprintf(ofp, "Query:       %s  [M=%d]\n", hmm->name, hmm->M);
    p7_ProfileConfig(hmm, bg, gm, 100, p7_LOCAL);
    p7_oprofile_Convert(gm, om);
    wom = p7_oprofile_Clone(om);
    pli = p7_pipeline_Create(NULL, om->M, 100, FALSE, p7_SEARCH_SEQS);
    if (F1 > 0.0) pli->F1 = F1;
    if (nobias) pli->do_biasfilter = FALSE;
    p7_pli_NewModel(pli, wom, bg);
    ox = p7_omx_Create(om->M, 0, 0);
    sq = esl_sq_CreateDigital(abc);

#ifdef MI355X_BLOCK
    ESL_SQ **blk = NULL;
    int      nblk = 0, capblk = 0, k = 0;
    if (workers > 0) {
      /* hmmsearch's threaded search: per-worker objects (src/hmmsearch.c:
       * 505-510), the read/search loop, then the merge (:538-542) */
      P7_MI355X_WORKER *info = calloc(workers, sizeof *info);
      const double      t0 = now();
      int               w;
      if (info == NULL) die("out of memory", NULL);
      for (w = 0; w < workers; w++) {
        info[w].om  = p7_oprofile_Clone(om);
        info[w].bg  = p7_bg_Clone(bg);
        info[w].th  = p7_tophits_Create();
        info[w].pli = p7_pipeline_Create(NULL, om->M, 100, FALSE, p7_SEARCH_SEQS);
        if (F1 > 0.0) info[w].pli->F1 = F1;
        if (nobias) info[w].pli->do_biasfilter = FALSE;
        p7_pli_NewModel(info[w].pli, info[w].om, info[w].bg);
      }
      sstatus = p7_mi355x_ThreadLoop(info, workers, sqfp, block_seqs, -1);
      if (sstatus == eslEFORMAT) die("sequence file parse failed:", esl_sqfile_GetErrorBuf(sqfp));
      if (sstatus != eslEOF) die("threaded block search failed:", info[0].pli->errbuf);
      for (w = 1; w < workers; w++) {
        p7_tophits_Merge(info[0].th, info[w].th);
        p7_pipeline_Merge(info[0].pli, info[w].pli);
      }
      fprintf(stderr, "search_seconds %s %.6f (%d workers, blocks of %d)\n", hmm->name, now() - t0, workers,
              block_seqs);
      p7_pipeline_Destroy(pli);
      p7_tophits_Destroy(th);
      pli = info[0].pli;
      th  = info[0].th;
      if (nolines) nseq = pli->nseqs;        /* no listing pass to count them */
      for (w = 0; w < workers; w++) {
        if (w > 0) { p7_pipeline_Destroy(info[w].pli); p7_tophits_Destroy(info[w].th); }
        p7_oprofile_Destroy(info[w].om);
        p7_bg_Destroy(info[w].bg);
      }
      free(info);
      /* the listing: a second pass over the file */
      if (!nolines && esl_sqfile_Position(sqfp, 0) != eslOK) die("cannot rewind", argv[2]);
    } else {
      /* the whole database as one block: read, searched, then listed */
      const double t0 = now();               /* reading included, as in the serial loop */
      for (;;) {
        ESL_SQ *x = esl_sq_CreateDigital(abc);
        if ((sstatus = esl_sqio_Read(sqfp, x)) != eslOK) { esl_sq_Destroy(x); break; }
        if (nblk == capblk) {
          capblk = 2 * capblk + 1024;
          if ((blk = realloc(blk, sizeof(ESL_SQ *) * capblk)) == NULL) die("out of memory", NULL);
        }
        blk[nblk++] = x;
      }
      if (sstatus != eslEOF) die("sequence file parse failed:", argv[2]);
      {
        const double t1 = now();
        if (p7_mi355x_PipelineBlock(pli, wom, bg, blk, nblk, th) != eslOK) die("block pipeline failed:", pli->errbuf);
        fprintf(stderr, "search_seconds %s %.6f (block step %.6f)\n", hmm->name, now() - t0, now() - t1);
      }
    }
    for (;;) {
      if (workers > 0) {
        if (nolines || (sstatus = esl_sqio_Read(sqfp, sq)) != eslOK) break;
      } else {
        if (k >= nblk) break;
        esl_sq_Copy(blk[k], sq);
        esl_sq_Destroy(blk[k++]);
        sstatus = eslOK;
      }
#else
    const double t0 = now();
    if (workers > 0) {
      /* hmmsearch --cpu N: per-worker objects (src/hmmsearch.c:505-510), the
       * threaded loop, the merge (:538-542); output without the listing */
      sse_worker *info = calloc(workers, sizeof *info);
      int         w;
      if (info == NULL) die("out of memory", NULL);
      if (!nolines) die("workers > 0 needs nolines in this build", NULL);
      for (w = 0; w < workers; w++) {
        info[w].om  = p7_oprofile_Clone(om);
        info[w].bg  = p7_bg_Clone(bg);
        info[w].th  = p7_tophits_Create();
        info[w].pli = p7_pipeline_Create(NULL, om->M, 100, FALSE, p7_SEARCH_SEQS);
        if (F1 > 0.0) info[w].pli->F1 = F1;
        if (nobias) info[w].pli->do_biasfilter = FALSE;
        p7_pli_NewModel(info[w].pli, info[w].om, info[w].bg);
      }
      sstatus = sse_threaded(info, workers, sqfp, abc, block_seqs);
      for (w = 1; w < workers; w++) {
        p7_tophits_Merge(info[0].th, info[w].th);
        p7_pipeline_Merge(info[0].pli, info[w].pli);
      }
      p7_pipeline_Destroy(pli);
      p7_tophits_Destroy(th);
      pli  = info[0].pli;
      th   = info[0].th;
      nseq = pli->nseqs;
      for (w = 0; w < workers; w++) {
        if (w > 0) { p7_pipeline_Destroy(info[w].pli); p7_tophits_Destroy(info[w].th); }
        p7_oprofile_Destroy(info[w].om);
        p7_bg_Destroy(info[w].bg);
      }
      free(info);
    }
    while (workers == 0 && (sstatus = esl_sqio_Read(sqfp, sq)) == eslOK) {
#endif
      float usc = 0.0f, vsc = 0.0f;
      int   st, vst;
      union { float f; uint32_t u; } bits;
      char  line[256];
      int   n;

      if (badlength) {                         /* om configured for another length than the target's */
        p7_oprofile_ReconfigLength(wom, 3 * sq->n + 100);
        if (badlength == 1) st = p7_MSVFilter(sq->dsq, (int) sq->n, wom, ox, &usc);
        else st = p7_ViterbiFilter(sq->dsq, (int) sq->n, wom, ox, &vsc);
        fprintf(stderr, "pipeline_driver: filter status %d on a mismatched length model\n", st);
        badlength = 0;
      }
#ifndef MI355X_BLOCK
      p7_pli_NewSeq(pli, sq);
      p7_bg_SetLength(bg, sq->n);
      p7_oprofile_ReconfigLength(wom, sq->n);
      p7_Pipeline(pli, wom, bg, sq, th);
#else
      p7_oprofile_ReconfigLength(wom, sq->n);
#endif
      nseq++;
      if (nolines) {
        esl_sq_Reuse(sq);
        p7_pipeline_Reuse(pli);
        continue;
      }

      st = p7_MSVFilter(sq->dsq, (int) sq->n, wom, ox, &usc);
      vst = p7_ViterbiFilter(sq->dsq, (int) sq->n, wom, ox, &vsc);
      bits.f = usc;
      n = snprintf(line, sizeof line, "msv %s %" PRId64 " %d %08x", sq->name, (int64_t) sq->n, st, bits.u);
      bits.f = vsc;
      if (n > 0 && n < (int) sizeof line)
        n += snprintf(line + n, sizeof line - n, " vit %d %08x\n", vst, bits.u);
      if (n < 0 || n >= (int) sizeof line) die("sequence name too long:", sq->name);
      if (nlines + n + 1 > cap) {
        cap = 2 * cap + 4096;
        if ((lines = realloc(lines, cap)) == NULL) die("out of memory", NULL);
      }
      memcpy(lines + nlines, line, n);
      nlines += n;
      esl_sq_Reuse(sq);
      p7_pipeline_Reuse(pli);
    }
#ifdef MI355X_BLOCK
    free(blk);
    if (workers > 0 && !nolines && sstatus != eslEOF) die("sequence file parse failed:", argv[2]);
#else
    if (sstatus != eslEOF) die("sequence file parse failed:", argv[2]);
    if (nolines) fprintf(stderr, "search_seconds %s %.6f\n", hmm->name, now() - t0);
#endif

    p7_tophits_SortBySortkey(th);
    p7_tophits_Threshold(th, pli);
    p7_tophits_Targets(ofp, th, pli, 120);  fprintf(ofp, "\n\n");
    p7_tophits_Domains(ofp, th, pli, 120);  fprintf(ofp, "\n\n");
    p7_tophits_TabularTargets(ofp, hmm->name, hmm->acc, th, pli, 1);
    p7_pli_Statistics(ofp, pli, NULL);
    fprintf(ofp, "//\n");
    if (nlines) fwrite(lines, 1, nlines, ofp);
    fprintf(ofp, "targets %" PRId64 "\n", nseq);

    free(lines);
    esl_sq_Destroy(sq);
    p7_omx_Destroy(ox);
    p7_pipeline_Destroy(pli);
    p7_tophits_Destroy(th);
    p7_oprofile_Destroy(wom);
    p7_oprofile_Destroy(om);
    p7_profile_Destroy(gm);
    p7_hmm_Destroy(hmm);
  }
  if (hstatus != eslEOF) die("HMM file parse failed:", argv[1]);

  fclose(ofp);
  if (sqfp) esl_sqfile_Close(sqfp);
  if (bg) p7_bg_Destroy(bg);
  if (abc) esl_alphabet_Destroy(abc);
  p7_hmmfile_Close(hfp);
  return 0;
}
