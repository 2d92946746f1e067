/*
 * hera_dropin.h -- the entry point hera's `quant` command calls.
 *
 *   EM_val *estimate_count(unsigned int *count, unsigned int order);
 *       replaces src/EM.c:315-387 (declared src/EM.h:28).  Reads hera's
 *       globals REF_INF (n, len, count) and CLASS (n, cls[i].{n, ref,
 *       count}) exactly as the reference does, runs the SQUAREM EM on the
 *       GPU with the reference's defaults (src/EM.c:324-346) and returns a
 *       malloc'd EM_val whose grama (TPM) and overlap (est_counts) are
 *       calloc'd, so hera's destroy_emVal / free release them.  `count` is
 *       the unique-read vector (REF_INF->count, or a bootstrap resample) and
 *       is not retained; `order` 0 is the main run, > 0 a bootstrap index,
 *       and selects the log line as in src/EM.c:375-384.  On a device error
 *       it prints "ERROR: ..." and exits, hera's convention
 *       (src/macros.h:6-9).
 *
 *   void destroy_emVal(EM_val *em_val);   src/EM.c:389-394
 *
 *   int hem_dropin_release(void);
 *       frees the device copy of the class table cached across calls (the
 *       table is fixed once alignment ends, so repeated calls - the main
 *       run and every bootstrap replicate - share one upload).
 *
 * REF_INF and CLASS are weak definitions in the library: hera's own
 * definitions (src/hash_align.c:35,42) take precedence when it links the
 * library.
 */
#ifndef HERA_DROPIN_H_
#define HERA_DROPIN_H_

#include "hera_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

extern Ref_inf *REF_INF;
extern Gclass *CLASS;

EM_val *estimate_count(unsigned int *count, unsigned int order);
void destroy_emVal(EM_val *em_val);
int hem_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
