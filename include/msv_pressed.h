/* msv_pressed.h -- reader for HMMER 3.1's pressed profile files.
 *
 * hmmpress writes, per model, the MSV part of P7_OPROFILE to <hmmfile>.h3f
 * and the rest to <hmmfile>.h3p (p7_oprofile_Write, src/impl_sse/io.c:
 * 89-180, format 3/f).  This reader replaces p7_oprofile_ReadMSV and the
 * ViterbiFilter part of p7_oprofile_ReadRest (io.c:204-300, :510-580) for
 * this library: it returns each model's fields de-striped into the
 * model-order tables msvg_profile_add and msvg_profile_set_viterbi take, so
 * a pressed database (hmmscan's input) loads without HMMER.
 */
#ifndef MSV_PRESSED_H
#define MSV_PRESSED_H

#include <stdint.h>
#include "msv_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msvh_pressed msvh_pressed;

/* Open a .h3f file and, optionally (h3p_path may be NULL), its .h3p
 * companion.  Returns 0, or 5 (eslEMEM) / 6 (eslENOTFOUND). */
int  msvh_pressed_open(const char *h3f_path, const char *h3p_path, msvh_pressed **ret);

/* Read the next model.  *msv is filled (rbv points into the reader's
 * buffer); *vit too when a .h3p is open (vit->M = 0 otherwise), with its
 * E-value parameters from the .h3f.  name receives the model name (NUL
 * terminated, truncated to name_cap).  The arrays stay valid until the next
 * call or msvh_pressed_close.  Returns 0, 3 (eslEOF) after the last model,
 * or 7 (eslEFORMAT) for a file that is not 3/f or is truncated or whose two
 * parts disagree. */
int  msvh_pressed_next(msvh_pressed *ph, msvg_oprofile *msv, msvg_vprofile *vit, char *name, int name_cap);

void msvh_pressed_close(msvh_pressed *ph);

#ifdef __cplusplus
}
#endif
#endif
