/* CPU restatements of the reference algorithms (test infrastructure only). */
#ifndef GTRAP_ORACLE_H
#define GTRAP_ORACLE_H
#include <stdint.h>

/* gtrap/gfilter.py:13-85 -- bisection median with zero halo; mode 0 sub, 1 div. */
void orc_gfilter_bisect(const float *in, float *out, int n, int nq, int r, int mode);

#endif

/* include/geebls.h:1-167 (simple=0) and gtrap/geebls_simple.py:17-175 (simple=1).
 * Bin sums are exact fixed-point sums with per-curve scale 2^-S (see bls_scale_exp),
 * rounded once to float32.  phase1/phase2 may be NULL. */
int orc_bls_scale_exp(float maxabs, uint32_t n);
void orc_geebls(float *blsp, uint32_t *phase1, uint32_t *phase2, const float *lc,
                const float *tu, const float *offset, const uint32_t *ntrue, uint32_t n,
                uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma, uint32_t kmi,
                uint32_t kkmi, float fmin, float df, uint32_t nthread, int simple,
                uint32_t q0, uint32_t q1, uint32_t f0, uint32_t f1);
