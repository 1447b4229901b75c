/* filters_mi355x.h -- the HMMER-side entry points of filters_mi355x.c, for
 * code compiled against HMMER 3.1b1's headers (hmmer.h).
 *
 * p7_MSVFilter and p7_ViterbiFilter keep their impl_sse.h declarations
 * (src/impl_sse/impl_sse.h:327, :343); this header adds the block step.
 */
#ifndef FILTERS_MI355X_H
#define FILTERS_MI355X_H

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

#endif
