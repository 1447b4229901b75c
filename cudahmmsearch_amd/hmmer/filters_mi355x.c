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
 * are not read: the library derives them from L with the expressions
 * p7_oprofile_ReconfigLength() uses (p7_oprofile.c:897-962).  This path
 * makes one host-device round trip per target; msvg_search and
 * msvg_vit_search are the batch paths.
 */
#include "p7_config.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "easel.h"
#include "hmmer.h"

#include "msv_gpu.h"

struct mi355x_state {
  msvg_ctx *ctx;
  int32_t   pid;      /* uploaded profile, or -1                              */
  uint8_t  *key;      /* the uploaded profile's fields, byte for byte         */
  size_t    keylen, keycap;
  uint8_t  *tmp;      /* the same for the profile of the current call         */
  size_t    tmpcap;
};

static __thread struct mi355x_state g_st = { NULL, -1, NULL, 0, 0, NULL, 0 };

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
  struct mi355x_state *s = &g_st;
  int status;

  if (p7O_NQB(om->M) > ox->allocQ16) ESL_EXCEPTION(eslEINVAL, "DP matrix allocated too small");  /* msvfilter.c:103 */
  ox->M = om->M;                                                                             /* :104 */
  if ((status = prepare(s, om)) != eslOK) return status;
  status = msvg_MSVFilter(s->ctx, s->pid, dsq, L, ret_sc);
  if (status != MSVG_OK && status != MSVG_ERANGE)
    ESL_EXCEPTION(status, "p7_MSVFilter: %s", msvg_last_error(s->ctx));
  return status;
}

int p7_ViterbiFilter(const ESL_DSQ *dsq, int L, const P7_OPROFILE *om, P7_OMX *ox, float *ret_sc)
{
  struct mi355x_state *s = &g_st;
  int status;

  if (p7O_NQW(om->M) > ox->allocQ8) ESL_EXCEPTION(eslEINVAL, "DP matrix allocated too small");  /* vitfilter.c:104 */
  if (om->mode != p7_LOCAL && om->mode != p7_UNILOCAL)                                          /* :105 */
    ESL_EXCEPTION(eslEINVAL, "Fast filter only works for local alignment");
  ox->M = om->M;                                                                                /* :106 */
  if ((status = prepare(s, om)) != eslOK) return status;
  status = msvg_ViterbiFilter(s->ctx, s->pid, dsq, L, ret_sc);
  if (status != MSVG_OK && status != MSVG_ERANGE)
    ESL_EXCEPTION(status, "p7_ViterbiFilter: %s", msvg_last_error(s->ctx));
  return status;
}
