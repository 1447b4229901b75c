/* msv_hmmfile.h -- reader for HMMER3 ASCII profile files (.hmm).
 *
 * Replaces p7_hmmfile_OpenE / p7_hmmfile_Read for this library
 * (src/p7_hmmfile.c:1181-1185, ASCII parser read_asc30hmm :1295-1630,
 * formats HMMER3/a to HMMER3/f): it returns each model's probabilities,
 * transitions and calibrated E-value parameters as P7_HMM holds them
 * (hmmer.h P7_HMM), so a text model converts with msvh_profile_convert_abc
 * and msvh_vit_convert_abc (include/msv_prep.h) into the profiles
 * msvg_profile_add and msvg_profile_set_viterbi take.  Binary (hmmpress'ed
 * .h3m), HMMER2 and gzip'ed files are not read (pressed databases have their
 * own reader, include/msv_pressed.h).
 */
#ifndef MSV_HMMFILE_H
#define MSV_HMMFILE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msvh_hmmfile msvh_hmmfile;

/* evparam indices, p7_MMU .. p7_FLAMBDA (hmmer.h p7_evparams_e) */
#define MSVH_MMU     0
#define MSVH_MLAMBDA 1
#define MSVH_VMU     2
#define MSVH_VLAMBDA 3
#define MSVH_FTAU    4
#define MSVH_FLAMBDA 5

typedef struct {
  int          M;             /* model length (LENG)                                       */
  int          alphatype;     /* MSVH_RNA, MSVH_DNA or MSVH_AMINO (ALPH)                    */
  int          K, Kp;         /* 4/18 or 20/29                                             */
  const float *mat;           /* (M+1) x K match emission probabilities; row 0 is zero     */
  const float *ins;           /* (M+1) x K insert emission probabilities                   */
  const float *t;             /* (M+1) x 7 transitions, MM MI MD IM II DM DD (p7H_*)       */
  const float *compo;         /* K: the COMPO line, or NULL when the file has none         */
  float        evparam[6];    /* the STATS lines (MSVH_MMU ...); 0 when the file has none  */
  int          has_stats;     /* 1 when all three STATS lines were present (p7H_STATS)     */
  int          max_length;    /* MAXL, or 0                                                */
  const char  *name;          /* NAME                                                      */
  const char  *acc;           /* ACC, or ""                                                */
  const char  *desc;          /* DESC, or ""                                               */
} msvh_hmm;

/* Open an HMMER3 ASCII file.  Returns 0, 6 (eslENOTFOUND) if it cannot be
 * opened, 7 (eslEFORMAT) if its first line is not an HMMER3/a..f tag, or
 * 5 (eslEMEM). */
int  msvh_hmmfile_open(const char *path, msvh_hmmfile **ret);

/* Read the next model into *hmm (its arrays and strings stay valid until the
 * next call or msvh_hmmfile_close).  Returns 0; 3 (eslEOF) after the last
 * model; 7 (eslEFORMAT) for a malformed model; 10 (eslEINCOMPAT) when a
 * model's alphabet differs from the first model's.  As in p7_hmmfile_Read, a
 * model that the end of the file cuts short also returns 3 (eslEOF);
 * msvh_hmmfile_error then says where it ended (it is "" after a normal
 * end). */
int  msvh_hmmfile_next(msvh_hmmfile *hfp, msvh_hmm *hmm);

/* A description of the last error, or "". */
const char *msvh_hmmfile_error(const msvh_hmmfile *hfp);

void msvh_hmmfile_close(msvh_hmmfile *hfp);

#ifdef __cplusplus
}
#endif
#endif
