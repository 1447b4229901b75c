/* msv_oracle.h -- CPU restatement of HMMER3's p7_MSVFilter composite.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Profile tables are row-major: rbv[x * M + (k-1)], x in 0..Kp-1, k in 1..M
 * (the de-striped form of P7_OPROFILE.rbv, impl_sse/p7_oprofile.c:625-630).
 * Targets are digital residue codes 0..Kp-1 WITHOUT sentinels.
 */
#ifndef MSV_ORACLE_H
#define MSV_ORACLE_H
#include <stdint.h>

#define ORC_eslOK        0
#define ORC_eslEINVAL   11
#define ORC_eslERANGE   16
#define ORC_eslENORESULT 19

typedef struct {
  int M, Kp;
  const uint8_t *rbv;         /* Kp * M                                   */
  int tbm_b, tec_b, base_b, bias_b;
  float scale_b;
} orc_profile;

int   orc_tjb(int L);
int   orc_ssv(const orc_profile *p, const uint8_t *x, int L, float *usc);
int   orc_msv_bytes(const orc_profile *p, const uint8_t *x, int L, float *usc);
int   orc_msv(const orc_profile *p, const uint8_t *x, int L, float *usc);
float orc_score(int xJ, int tjb, int base_b, float scale_b);
float orc_nullsc(int L);
int   orc_pass(float usc, int L, float mu, float lambda, double F1);
int   orc_msv_batch(const orc_profile *p, const uint8_t *res, const int64_t *off,
                    const int32_t *len, int n, float *usc, int32_t *status, int nthreads);

/* Checker for the GPU path's decomposition (see msv_oracle.c). */
int orc_jfree_max(const orc_profile *p, const uint8_t *x, int L);
int orc_decide(const orc_profile *p, int Rp, int L, int *xJ);

#endif
