/* CPU restatements of the reference algorithms (test infrastructure only). */
#ifndef GTRAP_ORACLE_H
#define GTRAP_ORACLE_H
#include <stdint.h>

/* gtrap/gfilter.py:13-85 -- bisection median with zero halo; mode 0 sub, 1 div. */
void orc_gfilter_bisect(const float *in, float *out, int n, int nq, int r, int mode);

#endif
