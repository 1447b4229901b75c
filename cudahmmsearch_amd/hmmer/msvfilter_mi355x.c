/* p7_MSVFilter() for HMMER 3.1b1's impl_sse backend, computed on an MI355X
 * through libmsvgpu.
 *
 * Replaces the function of src/impl_sse/msvfilter.c:78-213 (declared at
 * src/impl_sse/impl_sse.h:327) with the same signature, the same status
 * codes (eslOK; eslERANGE with *ret_sc = +inf; eslEINVAL through
 * ESL_EXCEPTION) and a bit-identical *ret_sc.  Every caller of p7_MSVFilter
 * (p7_Pipeline, p7_MSVMu, the drivers) takes it unchanged.
 *
 * Build: compile this file against the reference headers and the include/
 * directory of this package, link it ahead of the impl_sse objects with the
 * stock msvfilter.o's own symbol renamed
 *     objcopy --redefine-sym p7_MSVFilter=p7_MSVFilter_sse msvfilter.o
 * (the rest of that file, p7_MSVFilter_longtarget for nhmmer among it, stays
 * as it is), and link -lmsvgpu.
 *
 * Each host thread keeps its own context and its own uploaded profile, as
 * each hmmsearch worker keeps its own P7_OPROFILE clone.  The uploaded
 * profile is recognised by the contents of the MSV fields, not by the
 * pointer: between queries a freed P7_OPROFILE and its successor can share
 * an address.  tjb_b is not read: the library derives it from L with the
 * expression p7_oprofile_ReconfigLength() uses (p7_oprofile.c:921-925).
 * This path makes one host-device round trip per target; msvg_search is the
 * batch path.
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
  int32_t   pid;      /* uploaded profile, or -1                            */
  uint8_t  *key;      /* scalars + striped rbv rows of the uploaded profile */
  size_t    keylen, keycap;
  uint8_t  *tmp;      /* the same for the profile of the current call       */
  size_t    tmpcap;
};

static __thread struct mi355x_state g_st = { NULL, -1, NULL, 0, 0, NULL, 0 };

/* The bytes that decide the MSV result: M, Kp, the MSV scalars, the
 * calibrated MSV E-value parameters, and rbv's Kp x Q stripes of 16 bytes
 * (row stride allocQ16, so rows are taken one at a time). */
static int make_key(const P7_OPROFILE *om, uint8_t **buf, size_t *ret_len, size_t *cap)
{
  const int Q = p7O_NQB(om->M), Kp = om->abc->Kp;
  const size_t head = 2 * sizeof(int) + 4 + 3 * sizeof(float);
  const size_t len = head + (size_t) Kp * Q * 16;
  uint8_t *p;
  int x;

  if (len > *cap) {
    if ((p = realloc(*buf, len)) == NULL) return eslEMEM;
    *buf = p;
    *cap = len;
  }
  p = *buf;
  memcpy(p, &om->M, sizeof(int));                       p += sizeof(int);
  memcpy(p, &Kp, sizeof(int));                          p += sizeof(int);
  *p++ = om->tbm_b; *p++ = om->tec_b; *p++ = om->base_b; *p++ = om->bias_b;
  memcpy(p, &om->scale_b, sizeof(float));               p += sizeof(float);
  memcpy(p, &om->evparam[p7_MMU], sizeof(float));       p += sizeof(float);
  memcpy(p, &om->evparam[p7_MLAMBDA], sizeof(float));   p += sizeof(float);
  for (x = 0; x < Kp; x++) {
    memcpy(p, om->rbv[x], (size_t) Q * 16);
    p += (size_t) Q * 16;
  }
  *ret_len = len;
  return eslOK;
}

/* De-stripe rbv (byte z of vector q holds node k = q + z*Q + 1,
 * p7_oprofile.c:625-630) into the row-major [x][k-1] table of the ABI. */
static int upload(struct mi355x_state *s, const P7_OPROFILE *om)
{
  const int M = om->M, Q = p7O_NQB(M), Kp = om->abc->Kp;
  msvg_oprofile d;
  uint8_t *rbv;
  int x, q, z, status;

  if ((rbv = malloc((size_t) Kp * M)) == NULL) return eslEMEM;
  for (x = 0; x < Kp; x++) {
    const uint8_t *row = (const uint8_t *) om->rbv[x];
    for (q = 0; q < Q; q++)
      for (z = 0; z < 16; z++) {
        const int k = q + z * Q + 1;
        if (k <= M) rbv[(size_t) x * M + k - 1] = row[q * 16 + z];
      }
  }
  d.M = M;             d.Kp = Kp;            d.rbv = rbv;
  d.tbm_b = om->tbm_b; d.tec_b = om->tec_b;  d.base_b = om->base_b; d.bias_b = om->bias_b;
  d.scale_b = om->scale_b;
  d.mmu = om->evparam[p7_MMU];
  d.mlambda = om->evparam[p7_MLAMBDA];
  if (s->pid >= 0) msvg_profile_remove(s->ctx, s->pid);
  s->pid = -1;
  status = msvg_profile_add(s->ctx, &d, &s->pid);
  free(rbv);
  if (status != MSVG_OK) s->pid = -1;
  return status;
}

int p7_MSVFilter(const ESL_DSQ *dsq, int L, const P7_OPROFILE *om, P7_OMX *ox, float *ret_sc)
{
  struct mi355x_state *s = &g_st;
  size_t keylen = 0;
  int status;

  if (p7O_NQB(om->M) > ox->allocQ16) ESL_EXCEPTION(eslEINVAL, "DP matrix allocated too small");
  ox->M = om->M;

  if (s->ctx == NULL && msvg_init(0, &s->ctx) != MSVG_OK)
    ESL_EXCEPTION(eslEINVAL, "p7_MSVFilter: no MI355X device available");

  if ((status = make_key(om, &s->tmp, &keylen, &s->tmpcap)) != eslOK) ESL_EXCEPTION(status, "allocation failed");
  if (s->pid < 0 || keylen != s->keylen || memcmp(s->tmp, s->key, keylen) != 0) {
    uint8_t *k = s->key;
    size_t cap = s->keycap;
    if ((status = upload(s, om)) != MSVG_OK) ESL_EXCEPTION(status, "p7_MSVFilter: %s", msvg_last_error(s->ctx));
    s->key = s->tmp;  s->keycap = s->tmpcap;  s->keylen = keylen;
    s->tmp = k;       s->tmpcap = cap;
  }

  status = msvg_MSVFilter(s->ctx, s->pid, dsq, L, ret_sc);
  if (status != MSVG_OK && status != MSVG_ERANGE)
    ESL_EXCEPTION(status, "p7_MSVFilter: %s", msvg_last_error(s->ctx));
  return status;
}
