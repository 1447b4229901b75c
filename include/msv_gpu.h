/* msv_gpu.h -- C ABI of the MI355X MSV filter (HMMER3 p7_MSVFilter on gfx950).
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns an Easel-compatible status (easel/easel.h:95-114); none calls exit()
 * or lets a C++ exception cross the boundary.  Host buffers are owned by the
 * caller; device buffers by the library.  A context is bound to one GPU and
 * must not be used from two host threads at once (use one context per thread,
 * as HMMER gives each worker its own P7_OMX / P7_OPROFILE clone,
 * hmmsearch.c:429-439, cudaHmmsearch.cu:508).
 *
 * Reference interfaces replaced (paths in forestcheng/cudahmmsearch, code/):
 *   msvg_MSVFilter     int p7_MSVFilter(const ESL_DSQ *dsq, int L,
 *                        const P7_OPROFILE *om, P7_OMX *ox, float *ret_sc)
 *                        src/impl_sse/impl_sse.h:327, src/impl_sse/msvfilter.c:78
 *   msvg_search        the GPU block loop of cudaHmmsearch: loadCUDA()
 *                        src/GPUFunction.cu:332-492 (GPUFunction.h:72-76) feeding
 *                        usc into p7_PipelineShort() (src/p7_pipeline.c:887-918)
 *                        with the MSV pass rule of p7_Pipeline (p7_pipeline.c:667-677)
 *   msvg_profile_add   allocCUDAFixed()/loadOM() (GPUFunction.cu:260-313): the
 *                        MSV fields of P7_OPROFILE (impl_sse.h:79-86) to the GPU
 *   msvg_db_add        loadDBQuad() + the per-block H2D copies (GPUFunction.cu:
 *                        219-255, 420-457), here once, length-sorted and resident
 *   msvg_init/free     getGPUInfo()/freeMem() (GPUFunction.cu:494-520)
 */
#ifndef MSV_GPU_H
#define MSV_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the values of Easel's (easel/easel.h:95-114). */
#define MSVG_OK         0   /* eslOK      */
#define MSVG_EMEM       5   /* eslEMEM    */
#define MSVG_EINVAL    11   /* eslEINVAL  */
#define MSVG_ERANGE    16   /* eslERANGE: score overflowed, usc = +inf */
#define MSVG_ENORESULT 19   /* eslENORESULT */
#define MSVG_EDEVICE  100   /* HIP runtime error (no Easel counterpart) */

/* Search modes. */
#define MSVG_MODE_PLAIN    0  /* one launch per profile, targets in caller order, static schedule */
#define MSVG_MODE_BATCHED  1  /* length-sorted, dynamic persistent schedule, tuned tiles, hipGraph */
#define MSVG_FLAG_EAGER  0x10 /* or'ed into BATCHED: launch without the graph, with per-stage timing */

typedef struct msvg_ctx msvg_ctx;

/* The MSV part of P7_OPROFILE (impl_sse.h:79-86) plus the two calibrated
 * E-value parameters the pass rule reads (om->evparam[p7_MMU], [p7_MLAMBDA]).
 * rbv is de-striped: rbv[x * M + (k-1)] is the biased cost of residue x at
 * node k (impl_sse/p7_oprofile.c:625-630), x in 0..Kp-1. */
typedef struct {
  int32_t        M;
  int32_t        Kp;
  const uint8_t *rbv;
  uint8_t        tbm_b, tec_b, base_b, bias_b;
  float          scale_b;
  float          mmu, mlambda;
} msvg_oprofile;

/* Per-search statistics (kernel times from HIP events on the search stream). */
typedef struct {
  double  ms_fast;         /* sum of fast-stage launch durations */
  double  ms_slow;         /* exact byte-MSV stage */
  double  ms_finalize;
  double  ms_total;        /* first launch start to last launch end */
  int64_t cells;           /* sum over pairs of M * L */
  int64_t n_slow;          /* pairs that ran the exact byte MSV */
  int64_t n_jrow;          /* pairs that ran the J-aware f16 stage */
  int32_t n_fast_launches;
  int32_t dropin_path;     /* after msvg_MSVFilter: 1 its one-launch kernel, 2 the search path */
  /* Batched searches: the persistent launches of the batched pass (one per
   * lane-group size, longest-M group first) and, for eager searches, each
   * launch's duration and that of the J-aware launches after them. */
  int32_t n_batch;
  int32_t batch_lanes[4];  /* lanes per target of each batched launch */
  int64_t batch_cells[4];  /* sum of M * L over the pairs each one scores */
  double  batch_ms[4];     /* HIP events around each launch on the search stream (eager only, else 0) */
  double  ms_jrow;         /* from the last batched launch's end to the J-aware launches' join (eager only) */
} msvg_stats;

int         msvg_init(int device, msvg_ctx **ret_ctx);
void        msvg_free(msvg_ctx *ctx);
const char *msvg_last_error(const msvg_ctx *ctx);

/* Profiles: copied to the device; the caller's buffers may be freed after.
 * EINVAL if M < 1 or M > 16384, Kp outside 1..255, or bias_b + base_b >= 255
 * (a range no profile built by p7_oprofile_Convert reaches). */
int msvg_profile_add(msvg_ctx *ctx, const msvg_oprofile *om, int32_t *ret_id);
int msvg_profile_remove(msvg_ctx *ctx, int32_t id);

/* Replace a profile's calibrated MSV E-value parameters. */
int msvg_profile_set_evparams(msvg_ctx *ctx, int32_t id, float mmu, float mlambda);

/* p7_MSVMu (src/evalues.c:240-286) on the GPU: scores the 200 i.i.d. length-200
 * calibration targets of p7_Calibrate (RNG esl_randomness_CreateFast(42),
 * evalues.c:81) with this profile and fits the Gumbel location for lambda. */
int msvg_calibrate_mu(msvg_ctx *ctx, int32_t id, double lambda, float *ret_mu);

/* Target database: residues are digital codes 0..Kp-1 without sentinels,
 * target i at res[off[i] .. off[i]+len[i]).  Copied to the device once,
 * remapped to the codes present, padded and length-sorted. */
int msvg_db_add(msvg_ctx *ctx, const uint8_t *res, const int64_t *off, const int32_t *len,
                int64_t n, int32_t *ret_id);
int msvg_db_remove(msvg_ctx *ctx, int32_t id);
int64_t msvg_db_size(msvg_ctx *ctx, int32_t id);

/* Score every listed profile against every target.  Outputs are row-major
 * [np][n] in the caller's target order; any may be NULL:
 *   usc    MSV score in nats, bit-identical to p7_MSVFilter's *ret_sc
 *   status 0 (eslOK) or 16 (eslERANGE, usc = +inf)
 *   pass   1 if the target passes the F1 filter of p7_Pipeline (P <= F1). */
int msvg_search(msvg_ctx *ctx, const int32_t *profile_ids, int32_t np, int32_t db_id,
                double F1, int mode, float *usc, uint8_t *status, uint8_t *pass);

/* The same, leaving results on the device (for timing with inputs resident);
 * msvg_results copies the last device-resident results to the host. */
int msvg_search_device(msvg_ctx *ctx, const int32_t *profile_ids, int32_t np, int32_t db_id,
                       double F1, int mode);
int msvg_results(msvg_ctx *ctx, float *usc, uint8_t *status, uint8_t *pass);

/* Allocation-free batched search.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc).  The same search as
 * msvg_search(..., MSVG_MODE_BATCHED, ...) and bit for bit the same
 * results, written to the caller's device arrays d_usc, d_status, d_pass
 * ([np][n] in the caller's target order; any may be NULL).  Every table,
 * queue, counter and result word of the call lives in the library's static
 * device scratch array of 128 MiB per GPU (held until the process ends):
 * the call makes no hipMalloc, hipFree or other allocating call.  Limits:
 * EINVAL if the call needs more than the 128 MiB (about 2.6 bytes per
 * (profile, target) pair plus each profile's tables; the message gives the
 * need) or a profile takes the column-segmented path (M above the batched
 * kernel's LDS tables, about 4096); ENORESULT if more pairs than np*n/16
 * need the exact byte MSV.  All calls on one GPU share the scratch: calls
 * that overlap on one GPU (from several host threads or contexts) race.
 * Returns when the results are in place. */
int msvg_search_noalloc(msvg_ctx *ctx, const int32_t *profile_ids, int32_t np, int32_t db_id, double F1,
                        float *d_usc, uint8_t *d_status, uint8_t *d_pass);
int msvg_last_stats(msvg_ctx *ctx, msvg_stats *out);

/* Drop-in for p7_MSVFilter: dsq is 1-based with sentinels at 0 and L+1
 * (easel.h:308-309).  Returns MSVG_OK or MSVG_ERANGE; *ret_sc as the
 * reference sets it.  One kernel launch per call where the profile's table
 * over its whole alphabet fits LDS (M up to 2560 for amino, 4096 for DNA);
 * empty targets, longer profiles and the rare pairs needing the exact
 * sub-B0 maximum take the search path (msvg_search on a one-target
 * database), with the same results. */
int msvg_MSVFilter(msvg_ctx *ctx, int32_t profile_id, const uint8_t *dsq, int L, float *ret_sc);

/* hmmscan's direction: many profiles against one query (dsq 1-based with
 * sentinels, as for msvg_MSVFilter).  For profile i: usc[i], status[i]
 * and pass[i] (F1 rule of p7_pipeline.c:662-677 with its mu, lambda) as
 * msvg_search would give them for a one-target database — replaces the
 * per-model p7_MSVFilter calls that p7_Pipeline makes in hmmscan's loop over
 * a pressed database (src/hmmscan.c:1176-1190, serial_loop).  One launch per register class,
 * one workgroup per profile; the profiles' tables are built on the first
 * call and kept while the same profile list is scanned.  All profiles must
 * share one alphabet (EINVAL otherwise).  ms_fast in msvg_last_stats is the
 * kernels' time, cells M * L summed over the profiles. */
int msvg_scan(msvg_ctx *ctx, int32_t np, const int32_t *profile_ids, const uint8_t *dsq, int L, double F1,
              float *usc, uint8_t *status, uint8_t *pass);

/* msvg_scan for a batch of nq queries (digital codes, no sentinels: query q
 * is res[qoff[q] .. qoff[q] + qlen[q])): the result of profile i against
 * query q at usc[q * np + i] (status, pass likewise).  Each workgroup stages
 * its profile's table once and its waves share the queries, so a batch
 * costs far less than nq single calls.  No counterpart in the reference. */
int msvg_scan_many(msvg_ctx *ctx, int32_t np, const int32_t *profile_ids, int32_t nq, const uint8_t *res,
                   const int64_t *qoff, const int32_t *qlen, double F1, float *usc, uint8_t *status,
                   uint8_t *pass);

/* ---- The ViterbiFilter stage: the filter after MSV in p7_Pipeline
 * (p7_pipeline.c:699-705, the call at :701), int16 Viterbi of impl_sse/vitfilter.c:84-250. ----
 *
 * The ViterbiFilter fields of P7_OPROFILE (impl_sse.h:88-101), in model order
 * k = 1..M (the de-striped form of om->rwv / om->twv, p7_oprofile.c:653-740):
 *   rwv[x * M + (k-1)]  match emission of residue code x at node k
 *   twv[t * M + (k-1)]  t = 0..7 = p7O_BM, MM, IM, DM, MD, MI, II, DD: the value
 *                       the striped transition vector holds at node k's
 *                       position (BM, MM, IM, DM rotated by one: into k from
 *                       k-1; MD, MI, II, DD out of k), -32768 where the
 *                       reference pads.
 * xw[N|C|J][MOVE] are not passed: they follow from the target length as
 * p7_oprofile_ReconfigRestLength sets them (p7_oprofile.c:944-962, with nj);
 * the N/C/J loop costs are 0 as vf_conversion leaves them. */
typedef struct {
  int32_t M, Kp;
  const int16_t *rwv;
  const int16_t *twv;
  int16_t e_loop, e_move;      /* xw[E][LOOP], xw[E][MOVE] */
  int16_t base_w, ddbound_w;
  float scale_w, nj;
  float vmu, vlambda;          /* evparam[p7_VMU], evparam[p7_VLAMBDA] */
} msvg_vprofile;

/* Attach the ViterbiFilter fields to a profile added with msvg_profile_add
 * (same M and Kp).  Replaces p7_oprofile_Convert's vf_conversion output for
 * the device (p7_oprofile.c:653-740).  EINVAL if M, Kp differ or M > 16384. */
int msvg_profile_set_viterbi(msvg_ctx *ctx, int32_t profile_id, const msvg_vprofile *vp);

/* p7_ViterbiFilter on the listed targets (caller indices into the database):
 * sc[i] bit-identical to the reference's *ret_sc, status[i] 0 (eslOK) or 16
 * (eslERANGE, sc = +inf).  The MSV survivors of msvg_search (pass = 1) are
 * the list p7_Pipeline would score.  After the call msvg_last_stats holds the
 * kernel time in ms_fast and M * sum(L) in cells. */
int msvg_vit_search(msvg_ctx *ctx, int32_t profile_id, int32_t db_id, const int64_t *targets, int64_t n,
                    float *sc, uint8_t *status);

/* msvg_vit_search for np profiles at once over one database, profile i on
 * targets[offsets[i] .. offsets[i+1]) with its results at the same
 * positions of sc and status (p7_Pipeline's second filter for the MSV
 * survivors of a multi-profile search).  Each profile's launch lasts at
 * least as long as its longest target's row chain, so the launches go
 * round-robin over four streams of the context and those tails overlap;
 * results equal np separate msvg_vit_search calls.  ms_fast is the whole
 * stage.  No counterpart in the reference. */
int msvg_vit_search_many(msvg_ctx *ctx, int32_t np, const int32_t *profile_ids, int32_t db_id,
                         const int64_t *offsets, const int64_t *targets, float *sc, uint8_t *status);

/* Drop-in for p7_ViterbiFilter (impl_sse.h:343, vitfilter.c:84): dsq 1-based
 * with sentinels.  Returns MSVG_OK or MSVG_ERANGE. */
int msvg_ViterbiFilter(msvg_ctx *ctx, int32_t profile_id, const uint8_t *dsq, int L, float *ret_sc);

/* p7_ViterbiMu (src/evalues.c:313-352) on the GPU: the 200 i.i.d. length-200
 * targets that follow p7_MSVMu's in p7_Calibrate's stream
 * (esl_randomness_CreateFast(42)), scored with the ViterbiFilter, Gumbel
 * location fitted for lambda. */
int msvg_calibrate_vmu(msvg_ctx *ctx, int32_t profile_id, double lambda, float *ret_vmu);

/* Measurement helper, no counterpart in the reference: the LDS read
 * bandwidth this device sustains (bytes/s), from a conflict-free
 * ds_read_b128 stream on every CU; bench.py reports the MSV kernel's LDS
 * traffic against it. */
int msvg_probe_lds(msvg_ctx *ctx, double *ret_bytes_per_s);
/* Measurement: the J-free recurrence of the 64-lane layout (Q = 16, 2048
 * model positions per target, one target per wave, 16 waves per CU) over L
 * rows per wave (L a multiple of 16), once in the kernels' f16 form
 * (v_pk_add_f16 clamp + v_pk_maximum3_f16) and once in an unsigned 16-bit
 * form (v_pk_add_u16 + v_pk_sub_u16 clamp + v_pk_max_u16) on the same
 * synthetic costs and residues: each launch's time, the cells of one
 * launch, and whether the two forms' maxima agree on every target.  No
 * counterpart in the reference (its SSE filter has one form,
 * impl_sse/msvfilter.c:150-153). */
int msvg_probe_forms(msvg_ctx *ctx, int32_t L, double *ms_f16, double *ms_u16, int64_t *cells, int32_t *agree);

/* Test hook: after k more device allocations, fail the rest as out of
 * memory (k < 0: never).  The environment variable MSVG_TEST_FAIL_ALLOC sets
 * the initial value.  No counterpart in the reference. */
void msvg_test_alloc_budget(long k);

/* Tile choice for the batched mode: lanes per target (8, 16, 32, 64) and waves
 * per workgroup for profiles of length M; 0 restores the built-in choice.
 * The autotuner in the Python package persists its findings through this. */
int msvg_set_tile(msvg_ctx *ctx, int32_t M_lo, int32_t M_hi, int32_t lanes, int32_t waves);
/* ViterbiFilter shape for models with M_lo <= M <= M_hi, taking effect at the
 * next msvg_profile_set_viterbi: lanes per target 16 or 32 (several targets
 * per wave) or 64 (waves 1, 2, 4 or 8 per target).  lanes = 0 clears. */
int msvg_set_vit_tile(msvg_ctx *ctx, int32_t M_lo, int32_t M_hi, int32_t lanes, int32_t waves);
/* How the ViterbiFilter rows of profiles set after this call treat the
 * lazy-F test (vitfilter.c:199): -1 (default) by the profile's estimated
 * rate (msvh_vit_lazyf_rate: below 0.75 the rows vote on the test, else
 * every row closes the delete paths), 0 every row closes, 1 every row votes.
 * Scores and statuses are the same in all three. */
int msvg_set_vit_vote(msvg_ctx *ctx, int32_t mode);

/* Models too long for one wave's ViterbiFilter tables (the multi-wave
 * shapes) run as column segments of at most 1792 positions, one wave per
 * target, carrying each row's boundary values from one segment's launch to
 * the next (on = 1, the default), or on the multi-wave kernel (0), taking
 * effect at the next search.  A pair some row of which could raise J (or
 * overflow) is scored again by the multi-wave kernel, so the results never
 * depend on the setting; after a ViterbiFilter search n_slow counts those
 * pairs.  No counterpart in the reference. */
int msvg_set_vit_segments(msvg_ctx *ctx, int32_t on);
/* Tail guard (on by default): where a database's longest target would make
 * the tile's launch last longer than the whole database's work at the
 * device's rate (small databases holding a long target), the batched mode
 * takes the lane-group size that minimizes the larger of the two instead.
 * 0 keeps the tile table's choice everywhere.  No counterpart in the
 * reference. */
int msvg_set_tail_guard(msvg_ctx *ctx, int on);
/* The tile the batched mode uses for a profile of length M on a database:
 * lanes per target, waves per workgroup (0: derived from the LDS
 * footprint), and whether the model takes the column-segmented path. */
int msvg_tile_of(msvg_ctx *ctx, int32_t M, int32_t db_id, int32_t *lanes, int32_t *waves, int32_t *segmented);

/* ---- Serving: continuous batching over one resident database ----
 * The counterpart of the reference's hmmpgmd worker (src/hmmdwrkr.c:532-667),
 * which keeps a RAM-resident database (cachedb.h:19-34) and runs each search
 * command over it.  Here queries (profiles) arrive at any time; a server
 * thread admits each as soon as a slot, the HBM budget and the LDS budget
 * allow, and runs it on its own stream, so queries that arrive while others
 * are on the GPU start at once and their workgroups share the CUs.  While a
 * server runs on a context, make no other calls on that context. */
typedef struct msvg_server msvg_server;

typedef struct {
  int64_t ticket;
  int32_t status;          /* MSVG_OK, or the error the query ended with   */
  int32_t M;
  int64_t n_pass;          /* targets passing F1 (ERANGE included)          */
  int64_t cells;           /* M * residues of the database                  */
  double  t_submit, t_admit, t_done;   /* seconds since the server started */
  int32_t batch;           /* queries in the search that served this one    */
} msvg_query_result;

typedef struct {
  int64_t submitted, admitted, completed, pending, running;
  int64_t max_pending;
  double  busy_s;          /* time with at least one query on the GPU       */
  double  uptime_s;
  int64_t cached_profiles; /* query profiles kept on the device for reuse (at most 256; the least
                              recently used one that no running query holds is removed first) */
  int64_t batches;         /* searches started (admitted / batches: queries per search) */
} msvg_server_info;

/* slots: queries admitted to the GPU at once (1..16); hbm_budget: device
 * bytes the server may hold for running queries (0: no limit); F1 as in
 * msvg_search.  flags: MSVG_SERVE_PERSISTENT (each query's fast stage is one
 * persistent launch; an admitted query's workgroups start as the running
 * ones retire), MSVG_SERVE_SHARED (one chunk per workgroup, so running
 * queries' workgroups interleave on the CUs from the start) or
 * MSVG_SERVE_BATCHED (a free slot takes every waiting query, up to 16, into
 * one multi-profile search: msvg_search's batched mode, one persistent
 * launch per lane-group size over all of them; each query's result equals
 * the same query served alone).  hmmpgmd's worker runs one search command
 * per query (src/hmmdwrkr.c:603-622); batched serving folds the commands
 * that arrive together into one pass over the resident database. */
#define MSVG_SERVE_PERSISTENT 0
#define MSVG_SERVE_SHARED     1
#define MSVG_SERVE_BATCHED    2
int msvg_server_start(msvg_ctx *ctx, int32_t db_id, int32_t slots, double F1, int64_t hbm_budget,
                      int32_t flags, msvg_server **ret);
/* Queue a query; the profile is copied.  Thread-safe. */
int msvg_server_submit(msvg_server *srv, const msvg_oprofile *om, int64_t *ret_ticket);
/* Wait up to timeout_s (< 0: forever) for a query.  On MSVG_OK fills *out
 * and, if idx is not NULL, up to cap passing targets (caller order,
 * ascending) with their usc; the query's results are then released.
 * Returns MSVG_ENORESULT if the query has not finished in time. */
int msvg_server_wait(msvg_server *srv, int64_t ticket, double timeout_s, msvg_query_result *out, int32_t *idx,
                     float *usc, int64_t cap);
int msvg_server_stats(msvg_server *srv, msvg_server_info *out);
/* Stop admitting, finish the queries on the GPU, drop the queued ones. */
int msvg_server_stop(msvg_server *srv);

#ifdef __cplusplus
}
#endif
#endif
