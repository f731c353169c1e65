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

/* ------------------------------------------------------------------------ */
/* geebls: include/geebls.h:1-167 (simple: gtrap/geebls_simple.py:17-175).   */
/*                                                                          */
/* The reference accumulates bin sums with LDS float atomics in arbitrary    */
/* order (geebls.h:75-77); this restatement (and the GPU kernel) sums each   */
/* bin exactly in 64-bit fixed point with the per-curve scale 2^-S and       */
/* rounds once to float32, which is one of the values the reference's        */
/* unordered float sum approximates and is independent of order.            */
/* S = 62 - e - b where maxabs < 2^e (frexp) and n < 2^b: any sum of at most */
/* n quantised samples stays below 2^62 in magnitude.                       */
/* ------------------------------------------------------------------------ */
int orc_bls_scale_exp(float maxabs, uint32_t n) {
  int e = 0;
  if (maxabs > 0.0f) frexpf(maxabs, &e); /* maxabs < 2^e */
  else e = -100;
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) <= (uint64_t)n) ++b; /* n < 2^b */
  int S = 62 - e - b;
  if (S > 200) S = 200;
  return S;
}

static int64_t quant(float v, int S) { return (int64_t)llrint(ldexp((double)v, S)); }

void orc_geebls(float *blsp, uint32_t *phase1, uint32_t *phase2, const float *lc,
                const float *tu, const float *offset, const uint32_t *ntrue, uint32_t n,
                uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma, uint32_t kmi,
                uint32_t kkmi, float fmin, float df, uint32_t nthread, int simple,
                uint32_t q0, uint32_t q1, uint32_t jf0, uint32_t jf1) {
  const uint32_t kmi_e = kmi < 1 ? 1 : kmi;            /* geebls.h:26-28 */
  const uint32_t kkmi_e = kkmi < 5 ? 5 : kkmi;         /* geebls.h:31-33, MINBIN 5 */
  const float rnb = (float)nb;
  const uint32_t mcount = (uint32_t)((float)n / (float)nthread); /* geebls.h:57 */
  const uint32_t nbin_used = mcount * nthread < n ? mcount * nthread : n;
  int64_t *Y = (int64_t *)malloc(sizeof(int64_t) * (nb + kma));
  uint32_t *Cn = (uint32_t *)malloc(sizeof(uint32_t) * (nb + kma));
  float *yf = (float *)malloc(sizeof(float) * (nb + kma));
  for (uint32_t q = q0; q < q1; ++q) {
    const float off = simple ? 0.0f : offset[q];
    const float rntrue = simple ? (float)n : (float)ntrue[q];   /* geebls.h:23 */
    float maxabs = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
      const float v = lc[(size_t)i * nq + q] - off;
      const float a = fabsf(v);
      if (a > maxabs) maxabs = a;
    }
    const int S = orc_bls_scale_exp(maxabs, n);
    for (uint32_t jf = jf0; jf < jf1; ++jf) {
      const float f0 = fmaf(df, (float)jf, fmin);              /* geebls.h:35 */
      memset(Y, 0, sizeof(int64_t) * (nb + kma));
      memset(Cn, 0, sizeof(uint32_t) * (nb + kma));
      for (uint32_t i = 0; i < nbin_used; ++i) {               /* geebls.h:57-80 */
        const float tx = simple ? tu[i] : tu[(size_t)i * nq + q];
        if (!simple && !(tx >= 0.0f)) continue;                /* INVT 0.0 */
        float ph = tx * f0;
        ph = ph - (float)(int)ph;
        const float fj = rnb * ph;
        const uint32_t j = fj > 0.0f ? (uint32_t)fj : 0u; /* saturating float->uint */
        if (j >= nb) continue; /* lands in the wrap area and is overwritten (geebls.h:84-95) */
        Y[j] += quant(lc[(size_t)i * nq + q] - off, S);
        Cn[j] += 1;
      }
      for (uint32_t k = 0; k < kma; ++k) { Y[nb + k] = Y[k]; Cn[nb + k] = Cn[k]; }
      for (uint32_t j = 0; j < nb + kma; ++j) yf[j] = (float)ldexp((double)Y[j], -S);
      float power = 0.0f;
      uint32_t jn1 = 0xFFFFFFFFu, jn2 = 0xFFFFFFFFu;
      for (uint32_t i = 0; i < nb; ++i) {                       /* geebls.h:105-140 */
        float s = 0.0f;
        uint32_t kk = 0;
        for (uint32_t j = i, k = 0; j < i + kma; ++j) {
          k = k + 1;
          s = s + yf[j];
          kk = kk + Cn[j];
          if (k > kmi_e && kk > kkmi_e) {
            const float rn1 = (float)kk;
            const float pw = (s * s) / (rn1 * (rntrue - rn1));
            if (pw > power) { power = pw; jn1 = i; jn2 = j; }
          }
        }
      }
      blsp[jf + (size_t)nf * q] = sqrtf(power);
      if (phase1) phase1[jf + (size_t)nf * q] = jn1;
      if (phase2) phase2[jf + (size_t)nf * q] = jn2;
    }
  }
  free(Y); free(Cn); free(yf);
}
