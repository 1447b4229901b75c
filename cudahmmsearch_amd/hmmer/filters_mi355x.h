/* filters_mi355x.h -- the HMMER-side entry points of filters_mi355x.c, for
 * code compiled against HMMER 3.1b1's headers (hmmer.h).
 *
 * p7_MSVFilter and p7_ViterbiFilter keep their impl_sse.h declarations
 * (src/impl_sse/impl_sse.h:327, :343); this header adds the block step and
 * the threaded search loop around it.
 */
#ifndef FILTERS_MI355X_H
#define FILTERS_MI355X_H

#include "esl_sqio.h"
#include "hmmer.h"

/* The block step of cudaHmmsearch (src/GPUFunction.cu:360-490) for one
 * query: both filters scored on the GPU for the n sequences sq[0..n)
 * (p7_MSVFilter for all of them in one batched search, p7_ViterbiFilter for
 * the MSV survivors in one batched call), then, for each sequence in order,
 * p7_pli_NewSeq and, for the MSV survivors (the sequences p7_PipelineShort
 * does not drop at its first test), p7_bg_SetLength,
 * p7_oprofile_ReconfigLength and p7_PipelineShort (src/p7_pipeline.c:887)
 * with its MSV score, hits into th, and p7_pipeline_Reuse.  pli, om, bg and th as p7_Pipeline takes them;
 * the sequences are read, not reused.  Returns eslOK or the first status
 * p7_PipelineShort returns; GPU failures raise through ESL_EXCEPTION. */
int p7_mi355x_PipelineBlock(P7_PIPELINE *pli, P7_OPROFILE *om, P7_BG *bg, ESL_SQ **sq, int n, P7_TOPHITS *th);

/* The same for an Easel sequence block, as a worker of hmmsearch holds it
 * (src/hmmsearch.c:1395-1411: the loop over block->list that calls
 * p7_Pipeline), each sequence esl_sq_Reuse'd afterwards as that loop does. */
int p7_mi355x_PipelineSqBlock(P7_PIPELINE *pli, P7_OPROFILE *om, P7_BG *bg, ESL_SQ_BLOCK *block, P7_TOPHITS *th);

/* One search worker's own objects, as hmmsearch's WORKER_INFO
 * (src/hmmsearch.c:36-44) holds them: its pipeline, its clone of the
 * optimized profile, its background model and its hit list. */
typedef struct {
  P7_PIPELINE *pli;
  P7_OPROFILE *om;
  P7_BG       *bg;
  P7_TOPHITS  *th;
} P7_MI355X_WORKER;

/* hmmsearch's thread_loop + pipeline_thread (src/hmmsearch.c:1317-1425)
 * with the block step as the worker's body: the calling thread reads dbfp
 * in blocks of up to block_seqs sequences (esl_sqio_ReadBlock, at most
 * n_targetseqs in all, or every sequence when n_targetseqs < 0) into a
 * queue of 2 * nworkers + 1 blocks, and nworkers threads each take a block
 * and run p7_mi355x_PipelineSqBlock on it with info[w]'s objects, so the
 * GPU filters of one block, the host tail of another and the reading of
 * the next run at the same time.  Returns as thread_loop does: eslEOF when
 * the input was read to its end (or n_targetseqs were read), a parse status
 * (eslEFORMAT; esl_sqfile_GetErrorBuf has the message) otherwise, or the
 * first failure of a worker.  The caller merges the workers' hit lists and
 * pipelines afterwards (p7_tophits_Merge, p7_pipeline_Merge,
 * src/hmmsearch.c:538-542); sorted, their output equals the serial
 * search's. */
int p7_mi355x_ThreadLoop(P7_MI355X_WORKER *info, int nworkers, ESL_SQFILE *dbfp, int block_seqs, int n_targetseqs);

#endif
