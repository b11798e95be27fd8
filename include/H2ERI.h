/*
 * H2ERI.h -- drop-in umbrella header for callers of H2P-ERI.
 *
 * Replaces src/H2ERI.h of the reference (which includes CMS.h,
 * H2ERI_typedef.h, H2ERI_shell_operations.h, H2ERI_partition.h,
 * H2ERI_build_H2.h, H2ERI_build_Coulomb.h and H2ERI_build_exchange.h): all of
 * those declarations live in h2eri.h.  The reference's callers also call
 * SIMINT's global init/finalize (tests/test_J_input_mat.c:10,70); this
 * library's integral engine has no global state, so both are no-ops here.
 */
#ifndef H2ERI_AMD_COMPAT_H
#define H2ERI_AMD_COMPAT_H

#include <math.h>

#include "h2eri.h"

static inline void simint_init(void) {}
static inline void simint_finalize(void) {}

#endif /* H2ERI_AMD_COMPAT_H */
