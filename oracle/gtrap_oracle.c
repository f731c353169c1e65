/*
 * gtrap_oracle.c -- CPU restatements of the reference's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in gtrap_amd/ links, loads or calls this
 * file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker.  Each function cites the reference lines it
 * restates.  Built with -ffp-contract=off: every fused multiply-add is an
 * explicit fmaf()/fma() call, so the arithmetic is exactly what is written.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "gtrap_oracle.h"

/* ------------------------------------------------------------------------ */
/* gfilter: gtrap/gfilter.py:13-85.  For row i of curve q the window is       */
/* buff[i-r .. i+r] with FILVAL 0.0 outside [0, n) (:25-36).  min/max over the */
/* window (:40-45); bisection on estim = (min+max)/2 counting inf/sup/equal    */
/* (:47-71), stop when inf <= r+1 and sup <= r+1; select mxinf / estim /       */
/* minsup (:72-74); out = x - val (:76).  Mode 1 divides instead.             */
/* ------------------------------------------------------------------------ */
void orc_gfilter_bisect(const float *in, float *out, int n, int nq, int r, int mode) {
  const int w = 2 * r + 1;
  float *buf = (float *)malloc(sizeof(float) * (size_t)w);
  for (int q = 0; q < nq; ++q) {
    for (int i = 0; i < n; ++i) {
      for (int d = 0; d < w; ++d) {
        const int j = i - r + d;
        buf[d] = (j >= 0 && j < n) ? in[(size_t)j * nq + q] : 0.0f;
      }
      float mn = buf[0], mx = buf[0];
      for (int d = 0; d < w; ++d) {
        if (buf[d] < mn) mn = buf[d];
        if (buf[d] > mx) mx = buf[d];
      }
      int it = 0, inf = 0, sup = 0, equal = 0;
      float estim = 0.0f, mxinf = 0.0f, minsup = 0.0f;
      for (;;) {
        estim = (mn + mx) / 2.0f;
        inf = sup = equal = 0;
        mxinf = mn;
        minsup = mx;
        for (int d = 0; d < w; ++d) {
          const float v = buf[d];
          if (v < estim) {
            inf++;
            if (v > mxinf) mxinf = v;
          } else if (v > estim) {
            sup++;
            if (v < minsup) minsup = v;
          } else {
            equal++;
          }
        }
        if (inf <= r + 1 && sup <= r + 1) break;
        if (inf > sup) mx = mxinf;
        else mn = minsup;
        it++;
        if (it > 2 * r) break;
      }
      float val;
      if (inf >= r + 1) val = mxinf;
      else if (inf + equal >= r + 1) val = estim;
      else val = minsup;
      const float x = in[(size_t)i * nq + q];
      out[(size_t)i * nq + q] = mode ? x / val : x - val;
    }
  }
  free(buf);
}
