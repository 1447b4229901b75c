/* vit_oracle.h -- CPU restatement of HMMER3's p7_ViterbiFilter.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP Viterbi filter.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Profile arrays are in model order, k = 1..M (the de-striped form of the
 * P7_OPROFILE Viterbi vectors, impl_sse/p7_oprofile.c:653-740):
 *   rwv[x * M + (k-1)]   match emission of residue x at node k
 *   twv[t * M + (k-1)]   t = 0..7 = BM, MM, IM, DM, MD, MI, II, DD: the value the
 *                        striped vector holds at node k's position, i.e. the
 *                        transitions INTO k (BM, MM, IM, DM, from node k-1;
 *                        vf_conversion stores them rotated by one) and OUT of
 *                        k (MD, MI, II, DD); -32768 where vf_conversion pads.
 * Positions beyond M hold -32768 in every array (vf_conversion's padding).
 * Targets are digital residue codes WITHOUT sentinels.
 */
#ifndef VIT_ORACLE_H
#define VIT_ORACLE_H
#include <stdint.h>

typedef struct {
  int M, Kp;
  const int16_t *rwv;          /* Kp * M */
  const int16_t *twv;          /* 8 * M  */
  int e_loop, e_move;          /* xw[E][LOOP], xw[E][MOVE]                  */
  int base_w, ddbound_w;
  float scale_w;
  float nj;                    /* expected J uses: 1 multihit, 0 unihit    */
} orc_vprofile;

/* xw[N|C|J][MOVE] for target length L (p7_oprofile.c:944-962). */
int orc_vit_move(const orc_vprofile *p, int L);
/* p7_ViterbiFilter (impl_sse/vitfilter.c:84-250): eslOK or eslERANGE. */
int orc_vit(const orc_vprofile *p, const uint8_t *x, int L, float *sc);
int orc_vit_batch(const orc_vprofile *p, const uint8_t *res, const int64_t *off,
                  const int32_t *len, int n, float *sc, int32_t *status, int nthreads);

#endif
