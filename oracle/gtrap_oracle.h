/* CPU restatements of the reference algorithms (test infrastructure only). */
#ifndef GTRAP_ORACLE_H
#define GTRAP_ORACLE_H
#include <stdint.h>

/* gtrap/gfilter.py:13-85 -- bisection median with zero halo; mode 0 sub, 1 div. */
void orc_gfilter_bisect(const float *in, float *out, int n, int nq, int r, int mode);


/* include/geebls.h:1-167 (simple=0) and gtrap/geebls_simple.py:17-175 (simple=1).
 * Bin sums are exact fixed-point sums with per-curve scale 2^-S (see bls_scale_exp),
 * rounded once to float32.  phase1/phase2 may be NULL. */
int orc_bls_scale_exp(float maxabs, uint32_t n);
void orc_geebls(float *blsp, uint32_t *phase1, uint32_t *phase2, const float *lc,
                const float *tu, const float *offset, const uint32_t *ntrue, uint32_t n,
                uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma, uint32_t kmi,
                uint32_t kkmi, float fmin, float df, uint32_t nthread, int simple,
                uint32_t q0, uint32_t q1, uint32_t f0, uint32_t f1);

/* Periodic trapezoid least squares: the statistic of gtrap/tls.py:44-79
 * (get_chi2Hmax, divx=True, sigma=1) on time folded at each trial period,
 * best over (epoch, W, L) per period as tlsfit's nanargmax (tls.py:82-110).
 * Direct evaluation (every trial sums over its window samples). */
void orc_tlsfold(const float *t, const float *x, int n, const double *periods, int np,
                 const double *widths, int nw, int nl, double dtep, int stat, double *best_sn,
                 int *best_w, int *best_l, int *best_e, double *best_h);
/* One trial (period P, epoch e, width W, L index m); returns S/N, writes H,
 * n, the residual res and the window mean (each output pointer may be null). */
double orc_tlsfold_trial(const float *t, const float *x, int n, double P, int e, double W,
                         int m, int nl, double dtep, int divx, int stat, double *H,
                         int *cnt, double *res, double *mean);

/* include/gtls_simple.h (mode 0) and include/gtls_absolute.h (mode 1), float32
 * line-faithful; outputs at k*nq + q for epochs k in [k0, k1), all curves. */
void orc_gtls(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
              const float *lc, const float *tu, int nt, int nq, int nw, int nl, int nsc,
              float wmax, float dw, float deltat, int mode, int k0, int k1);

/* gtls with float32 region membership (as include/gtls_simple.h decides it)
 * and float64 sums, H, residual and S/N on the exact shifted time (the build's
 * fast gtls mode); sn64 (may be null) receives the float64 best S/N. */
void orc_gtls64(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
                double *sn64, const float *lc, const float *tu, int nt, int nq, int nw, int nl,
                int nsc, float wmax, float dw, float deltat, int mode, int k0, int k1);

#endif /* GTRAP_ORACLE_H */
