/* vit_oracle.c -- CPU restatement of HMMER 3.1b1's p7_ViterbiFilter.
 *
 * TEST INFRASTRUCTURE ONLY (see vit_oracle.h).
 *
 * Follows impl_sse/vitfilter.c:84-250 one cell at a time, in model order
 * instead of Farrar's striped order.  The striped code applies the same
 * per-cell operations in the same order (every operation is elementwise),
 * so the order of cells does not change any value; two places need care:
 *
 *  - Padding.  The striped rows run over 8*Q >= M positions with -32768 in
 *    every padded emission and transition (p7_oprofile.c:683-714).  Padded
 *    match cells are exactly -32768 (every candidate is <= -1 before the
 *    -32768 emission saturates it), but the delete candidate one past the
 *    end, D(M+1) = M(M) + tMD(M) with tMD(M) = -32768, is M(M) - 32768 and
 *    can be above -32768.  It enters the row's Dmax (vitfilter.c:168,
 *    the max runs over every delayed M->D value), so positions 1..M+1 are
 *    computed here.
 *  - The lazy-F delete pass (vitfilter.c:183-232).  When it runs, the
 *    striped passes repeat until no lane improves, which is the closure
 *    D(k) = max(D(k), D(k-1) + tDD(k-1)) in model order; when it does not
 *    run, D keeps the M->D values only.  The decision is the reference's:
 *    Dmax + ddbound_w > xB.
 *
 * Saturating 16-bit adds (_mm_adds_epi16) are adds16; the special states
 * are C int16_t variables updated with int arithmetic (vitfilter.c:
 * 176-181).  N/C/J loop costs are 0 (vf_conversion, p7_oprofile.c:726-731).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vit_oracle.h"

#define ERANGE_ 16

static inline int adds16(int a, int b)
{
  int s = a + b;
  return s > 32767 ? 32767 : (s < -32768 ? -32768 : s);
}
static inline int imax(int a, int b) { return a > b ? a : b; }

/* wordify(), p7_oprofile.c:527-534 */
static int wordify(float scale_w, float sc)
{
  sc = roundf(scale_w * sc);
  if (sc >= 32767.0) return 32767;
  else if (sc <= -32768.0) return -32768;
  else return (int16_t) sc;
}

/* p7_oprofile_ReconfigRestLength, p7_oprofile.c:948-956 */
int orc_vit_move(const orc_vprofile *p, int L)
{
  float pmove = (2.0f + p->nj) / ((float) L + 2.0f + p->nj);
  return wordify(p->scale_w, logf(pmove));
}

int orc_vit(const orc_vprofile *p, const uint8_t *x, int L, float *sc)
{
  const int M = p->M, P = M + 1;
  const int16_t *T = p->twv;
  int *buf = malloc(sizeof(int) * 6 * (P + 1));
  int *Mp = buf, *Ip = Mp + (P + 1), *Dp = Ip + (P + 1);
  int *Mc = Dp + (P + 1), *Ic = Mc + (P + 1), *Dc = Ic + (P + 1), *tmp;
  const int move = orc_vit_move(p, L);
  int xN, xB, xJ, xC, xE, i, k;

#define TR(t, k) ((k) <= M ? (int) T[(t) * M + (k) - 1] : -32768)

  for (k = 0; k <= P; k++) Mp[k] = Ip[k] = Dp[k] = -32768;          /* vitfilter.c:115 */
  xN = p->base_w;                                                    /* :116 */
  xB = (int16_t) (xN + move);                                        /* :117 */
  xJ = -32768;
  xC = -32768;

  for (i = 0; i < L; i++) {
    const int16_t *e = p->rwv + (size_t) x[i] * M;
    int xEm = -32768, Dmax = -32768, dcv = -32768;
    for (k = 1; k <= P; k++) {
      const int mpv = (k > 1) ? Mp[k - 1] : -32768;                  /* :139-141 */
      const int ipv = (k > 1) ? Ip[k - 1] : -32768;
      const int dpv = (k > 1) ? Dp[k - 1] : -32768;
      int sv = adds16(xB, TR(0, k));                                 /* :146-150 */
      sv = imax(sv, adds16(mpv, TR(1, k)));
      sv = imax(sv, adds16(ipv, TR(2, k)));
      sv = imax(sv, adds16(dpv, TR(3, k)));
      sv = adds16(sv, k <= M ? e[k - 1] : -32768);
      xEm = imax(xEm, sv);                                           /* :151 */
      Mc[k] = sv;                                                    /* :161 */
      Dc[k] = dcv;                                                   /* :162 */
      dcv = adds16(sv, TR(4, k));                                    /* :167 */
      Dmax = imax(Dmax, dcv);                                        /* :168 */
      Ic[k] = imax(adds16(Mp[k], TR(5, k)), adds16(Ip[k], TR(6, k)));  /* :171-172 */
    }

    xE = xEm;                                                        /* :176 */
    if (xE >= 32767) { free(buf); *sc = INFINITY; return ERANGE_; }  /* :177 */
    xN = (int16_t) (xN + 0);                                         /* :178 */
    xC = (int16_t) imax(xC + 0, xE + p->e_move);                     /* :179 */
    xJ = (int16_t) imax(xJ + 0, xE + p->e_loop);                     /* :180 */
    xB = (int16_t) imax(xJ + move, xN + move);                       /* :181 */

    if (Dmax + p->ddbound_w > xB)                                    /* :199-232 */
      for (k = 2; k <= P; k++) Dc[k] = imax(Dc[k], adds16(Dc[k - 1], TR(7, k - 1)));

    tmp = Mp; Mp = Mc; Mc = tmp;
    tmp = Ip; Ip = Ic; Ic = tmp;
    tmp = Dp; Dp = Dc; Dc = tmp;
  }
#undef TR
  free(buf);

  if (xC > -32768) {                                                 /* :240-248 */
    float s = (float) xC + (float) move - (float) p->base_w;
    s /= p->scale_w;
    s -= 3.0;
    *sc = s;
  } else {
    *sc = -INFINITY;
  }
  return 0;
}

typedef struct {
  const orc_vprofile *p;
  const uint8_t *res;
  const int64_t *off;
  const int32_t *len;
  float *sc;
  int32_t *status;
  int n, next;
  pthread_mutex_t lock;
} vbatch;

static void *vworker(void *arg)
{
  vbatch *b = arg;
  for (;;) {
    int i0, i1, i;
    pthread_mutex_lock(&b->lock);
    i0 = b->next; b->next += 16; if (b->next > b->n) b->next = b->n;
    i1 = b->next;
    pthread_mutex_unlock(&b->lock);
    if (i0 >= i1) break;
    for (i = i0; i < i1; i++) b->status[i] = orc_vit(b->p, b->res + b->off[i], b->len[i], &b->sc[i]);
  }
  return NULL;
}

int orc_vit_batch(const orc_vprofile *p, const uint8_t *res, const int64_t *off,
                  const int32_t *len, int n, float *sc, int32_t *status, int nthreads)
{
  vbatch b;
  pthread_t *tid;
  int t;
  if (nthreads < 1) nthreads = 1;
  b.p = p; b.res = res; b.off = off; b.len = len; b.sc = sc; b.status = status;
  b.n = n; b.next = 0;
  pthread_mutex_init(&b.lock, NULL);
  tid = malloc(sizeof(pthread_t) * nthreads);
  for (t = 0; t < nthreads; t++) pthread_create(&tid[t], NULL, vworker, &b);
  for (t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
  free(tid);
  pthread_mutex_destroy(&b.lock);
  return 0;
}
