/* msv_prep.h -- host-side profile preparation and workload generation.
 *
 * The half of HMMER that sits in front of p7_MSVFilter: everything needed to
 * turn an HMM into the byte profile the filter reads, to compute the
 * length-dependent constants, and to regenerate the sampled workloads
 * (sampled HMMs, i.i.d. background targets, planted core emissions) on any
 * host.  Plain C ABI; all functions are pure host code and thread-safe for
 * distinct RNG objects.
 *
 * Reference functions mirrored (code/ in forestcheng/cudahmmsearch):
 *   msvh_rng_init      esl_randomness_Create / _CreateFast   easel/esl_random.c:84-145, :197-214
 *   msvh_random        esl_random                          easel/esl_random.c:271-276
 *   msvh_fchoose       esl_rnd_FChoose                     easel/esl_random.c:718-738
 *   msvh_hmm_sample    p7_hmm_Sample                       src/p7_hmm.c:877-918
 *   msvh_profile_convert  p7_ProfileConfig match emissions src/modelconfig.c:138-150 +
 *                      the MSV part of p7_oprofile_Convert  src/impl_sse/p7_oprofile.c:602-640
 *   msvh_lambda        p7_Lambda                           src/evalues.c:186-192
 *   msvh_core_emit     p7_CoreEmit                         src/emit.c:71-148
 *   msvh_iid           esl_rsq_xfIID over p7_AminoFrequencies  easel/esl_randomseq.c:764-772,
 *                                                          src/hmmer.c:163-186
 *   msvh_tjb           p7_oprofile_ReconfigMSVLength        src/impl_sse/p7_oprofile.c:921-925
 *   msvh_nullsc        p7_bg_SetLength + p7_bg_NullOne      src/p7_bg.c:191-198, :359-363
 *   msvh_fit_mu        esl_gumbel_FitCompleteLoc over the p7_MSVMu scores  easel/esl_gumbel.c:492-503,
 *                                                          src/evalues.c:240-286
 */
#ifndef MSV_PREP_H
#define MSV_PREP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSVH_K  20   /* amino-acid alphabet size */
#define MSVH_KP 29   /* with gap, degeneracies, nonresidue, missing data */

/* Alphabet types, Easel's codes (easel/esl_alphabet.h:12-14): amino has
 * K = 20, Kp = 29 ("ACDEFGHIKLMNPQRSTVWY-BJZOUX*~"); DNA and RNA K = 4,
 * Kp = 18 ("ACGT-RYMKSWHBVDN*~", U for T). */
#define MSVH_RNA   1
#define MSVH_DNA   2
#define MSVH_AMINO 3

typedef struct {
  int      fast;       /* 0: Mersenne twister (esl_randomness_Create), 1: LCG (CreateFast) */
  uint32_t mt[624];
  int      mti;
  uint32_t x;
  uint32_t seed;
} msvh_rng;

void   msvh_rng_init(msvh_rng *r, uint32_t seed, int fast);
double msvh_random(msvh_rng *r);
int    msvh_fchoose(msvh_rng *r, const float *p, int N);
void   msvh_background(float *f);             /* p7_AminoFrequencies, 20 floats */

/* mat, ins: (M+1) x 20 (row 0 of mat unused); t: (M+1) x 7 in
 * MM, MI, MD, IM, II, DM, DD order (p7H_* in hmmer.h). */
int    msvh_hmm_sample(msvh_rng *r, int M, float *mat, float *ins, float *t);

/* Byte profile from HMM match emissions: rbv is 29 x M (de-striped);
 * scalars = {tbm_b, tec_b, tjb_b (for config_L), base_b, bias_b};
 * *scale = scale_b.  msc (optional, 29 x M) receives the float match scores. */
int    msvh_profile_convert(int M, const float *mat, int config_L, uint8_t *rbv, int *scalars, float *scale,
                            float *msc);
double msvh_lambda(int M, const float *mat);
/* The same for any of the alphabets above: mat is (M+1) x K, rbv Kp x M,
 * msc Kp x M; the background is p7_bg_Create's (p7_bg.c:55-82: the amino
 * frequencies, uniform 1/K for DNA and RNA).  11 (eslEINVAL) for an unknown
 * alphabet or M < 1. */
int    msvh_profile_convert_abc(int alphatype, int M, const float *mat, int config_L, uint8_t *rbv, int *scalars,
                                float *scale, float *msc);
double msvh_lambda_abc(int alphatype, int M, const float *mat);

/* One core emission (residue codes, no sentinels) into out[0..cap); returns
 * its length, or -1 if cap is too small. */
int    msvh_core_emit(msvh_rng *r, int M, const float *mat, const float *ins, const float *t, uint8_t *out,
                      int cap);
void   msvh_iid(msvh_rng *r, int L, uint8_t *out);

int    msvh_tjb(int L);
float  msvh_nullsc(int L);

/* mu from N MSV scores (usc, status) of i.i.d. targets of length L as
 * p7_MSVMu computes it; ERANGE scores count as (255 - base_b) / scale_b. */
double msvh_fit_mu(const float *usc, const int32_t *status, int N, int L, int base_b, float scale_b, double lambda);

/* The targets p7_MSVMu scores: N i.i.d. sequences of length L from the
 * calibration RNG esl_randomness_CreateFast(42) (evalues.c:81, :261). */
void   msvh_calibration_targets(int N, int L, uint8_t *out);

/* ---- ViterbiFilter stage ---- */
/* The ViterbiFilter fields of p7_oprofile_Convert (vf_conversion,
 * p7_oprofile.c:653-740) for an HMM configured as hmmsearch configures it
 * (p7_ProfileConfig local multihit, modelconfig.c:49-150): rwv (29 x M) and
 * twv (8 x M) in model order as msvg_vprofile holds them; scalars =
 * {xw[E][LOOP], xw[E][MOVE], base_w, ddbound_w}; scale_w; nj.  mat and t as
 * msvh_hmm_sample writes them. */
int    msvh_vit_convert(int M, const float *mat, const float *t, int16_t *rwv, int16_t *twv, int *scalars,
                        float *scale_w, float *nj);
/* The same for any alphabet: mat (M+1) x K, rwv Kp x M. */
int    msvh_vit_convert_abc(int alphatype, int M, const float *mat, const float *t, int16_t *rwv, int16_t *twv,
                            int *scalars, float *scale_w, float *nj);
/* xw[N|C|J][MOVE] for target length L: wordify(logf((2+nj)/(L+2+nj)))
 * (p7_oprofile_ReconfigRestLength, p7_oprofile.c:944-962; wordify :527-534). */
int    msvh_vit_move(float scale_w, float nj, int L);
/* p7_ViterbiFilter's score from its final xC (vitfilter.c:240-248): -inf when
 * xC is -32768. */
float  msvh_vit_score(int xC, int move, int base_w, float scale_w);
/* Fraction of rows on which p7_ViterbiFilter's lazy-F test (vitfilter.c:199)
 * holds, on one pseudo-random target of `rows` residues: a planning estimate
 * (the kernels vote on the test per row when it is rare, and close the
 * delete paths on every row when it is not; scores are the same). */
double msvh_vit_lazyf_rate(const int16_t *rwv, const int16_t *twv, int M, int Kp, int e_loop, int base_w,
                           int ddbound_w, int move, int rows);
/* mu from N ViterbiFilter scores of i.i.d. targets of length L as
 * p7_ViterbiMu computes it (evalues.c:313-352); ERANGE scores count as
 * (32767 - base_w) / scale_w. */
double msvh_fit_vmu(const float *sc, const int32_t *status, int N, int L, int base_w, float scale_w, double lambda);
/* The targets p7_ViterbiMu scores in p7_Calibrate: the N sequences of length
 * L that follow p7_MSVMu's N in the same stream (evalues.c:107-108). */
void   msvh_vit_calibration_targets(int N, int L, uint8_t *out);

/* Synthetic target database: n targets, lengths log-normal (median, sigma)
 * clipped to [lo, hi], optionally a Pareto tail (fraction tail_frac, x_m =
 * tail_xm, alpha tail_alpha, clipped to tail_max), residues i.i.d. from the
 * amino background, generated in blocks of 4096 targets on nthreads threads
 * (block b uses Mersenne seed seed * 1000003 + b + 1).  lens must hold n;
 * call once with res = NULL to get the total residue count. */
int64_t msvh_gen_lengths(uint32_t seed, int64_t n, double median, double sigma, int lo, int hi,
                         double tail_frac, double tail_xm, double tail_alpha, int tail_max, int32_t *lens);
void   msvh_gen_residues(uint32_t seed, int64_t n, const int32_t *lens, const int64_t *off, uint8_t *res,
                         int nthreads);

#ifdef __cplusplus
}
#endif
#endif
