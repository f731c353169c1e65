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

/* ------------------------------------------------------------------------ */
/* Periodic trapezoid least squares (the fold extension of gtrap/tls.py).   */
/*                                                                          */
/* Folding: u = fmod(t, P) in float64 (exact); epoch e has t0 = e*dtep,      */
/* e < ceil(P/dtep).  Offset of a sample: d = u - t0, replaced by            */
/* (u - P) - t0 when d >= P/2 and by (u + P) - t0 when d < -P/2 (each        */
/* operation rounded in float64).  Masked samples (t < 0) and samples with   */
/* a non-finite flux are skipped.                                           */
/* Trial (W, L = m * (0.45*W/nl)): regions of tls.py:46-50 on d,             */
/* A [-W/2, -W/4-L/2), B [-W/4-L/2, -W/4+L/2), C [-W/4+L/2, W/4-L/2),        */
/* D [W/4-L/2, W/4+L/2), E [W/4+L/2, W/2); mean over [-W/2, W/2] (:57-59);   */
/* denA, numB, H = -numB/denA (:71-73); res (:74-77);                       */
/* S/N = -H / sqrt(res / (n - 3)) (tls.py:100); a trial with n <= 3 (no      */
/* degrees of freedom) is NaN, and so is one with W >= P (a window as long  */
/* as the period would hold two images of a sample).                        */
/* ------------------------------------------------------------------------ */
/* The value of a trial.  stat 0: tls.py's S/N = -H/sqrt(res/(n-3))
 * (tls.py:100).  stat 1: the least-squares fit of the trapezoid shape g
 * (tls.py:2-15 with unit depth), from a = sum g^2 and b = sum y g over the
 * window: H* = b/a and the signed chi^2 improvement sign(-H*) b^2/a, which
 * ranks the trials (the caller divides its root by the curve's standard
 * deviation to report it). */
static double trial_value(double H, double res, double a, double b, int n, int stat,
                          double *Hls) {
  if (!stat) return -H / sqrt(res / (double)(n - 3));
  if (!(a > 0.0)) return NAN;
  *Hls = b / a;
  const double k = b * b / a;
  return b > 0.0 ? -k : k;
}

/* root of a signed chi^2 improvement over sigma (stat 1's reported value) */
static double dchi2_value(double key, double sigma) {
  return (key >= 0.0 ? sqrt(key) : -sqrt(-key)) / sigma;
}

/* standard deviation of the usable samples' flux (sum of squares form, the
 * kernel's) */
static double flux_sigma(const float *t, const float *x, int n) {
  double sx = 0.0, sxx = 0.0;
  int c = 0;
  for (int i = 0; i < n; ++i) {
    if (!(t[i] >= 0.0f && isfinite(x[i]))) continue;
    sx += (double)x[i];
    sxx += (double)x[i] * (double)x[i];
    c++;
  }
  if (c <= 1) return NAN;
  const double v = sxx - sx * sx / (double)c;
  return sqrt((v > 0.0 ? v : 0.0) / (double)(c - 1));
}

/* a sample takes part when its time is valid (t >= 0) and its flux finite */
static int usable(float t, float x) { return t >= 0.0f && isfinite(x); }

static double fold_offset(double u, double t0, double P) {
  double d = u - t0;
  if (d >= 0.5 * P) d = (u - P) - t0;
  else if (d < -0.5 * P) d = (u + P) - t0;
  return d;
}

double orc_tlsfold_trial(const float *t, const float *x, int n, double P, int e, double W,
                         int m, int nl, double dtep, int divx, int stat, double *Hout,
                         int *cnt, double *resout, double *meanout) {
  const double t0 = (double)e * dtep;
  const double dl = 0.45 * W / (double)nl;
  const double L = (double)m * dl;
  const double hw = 0.5 * W, qw = 0.25 * W, hl = 0.5 * L;
  const double tAB = -qw - hl, tBC = -qw + hl, tCD = qw - hl, tDE = qw + hl;
  double msum = 0.0;
  int mcnt = 0;
  for (int i = 0; i < n; ++i) {
    if (!usable(t[i], x[i])) continue;
    const double d = fold_offset(fmod((double)t[i], P), t0, P);
    if (d >= -hw && d <= hw) { msum += (double)x[i]; mcnt++; }
  }
  const double mean = divx ? msum / (double)mcnt : 0.0; /* divx=False: tls.py:60-62 */
  int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
  double xAEC = 0, xBD = 0, xtB = 0, xtD = 0, tB = 0, tD = 0, t2BD = 0;
  for (int i = 0; i < n; ++i) {
    if (!usable(t[i], x[i])) continue;
    const double d = fold_offset(fmod((double)t[i], P), t0, P);
    const double y = (double)x[i] - mean;
    if (d >= -hw && d < tAB) { nA++; xAEC += y; }
    else if (d >= tAB && d < tBC) { nB++; xBD += y; xtB += y * d; tB += d; t2BD += d * d; }
    else if (d >= tBC && d < tCD) { nC++; xAEC += y; }
    else if (d >= tCD && d < tDE) { nD++; xBD += y; xtD += y * d; tD += d; t2BD += d * d; }
    else if (d >= tDE && d < hw) { nE++; xAEC += y; }
  }
  const double denA = 4.0 * L * L * (double)(nA - nC + nE) + W * W * (double)(nB + nD) +
                      8.0 * W * (tB - tD) + 16.0 * t2BD;
  const double numB = 8.0 * L * L * xAEC - 4.0 * W * L * xBD + 16.0 * L * (xtD - xtB);
  const double H = -numB / denA;
  const double h = H / 2.0;
  double res = 0.0, ga = 0.0, gb = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!usable(t[i], x[i])) continue;
    const double d = fold_offset(fmod((double)t[i], P), t0, P);
    const double y = (double)x[i] - mean;
    double r, gv;  /* gv: the trapezoid of unit depth, tls.py:2-15 */
    if (d >= -hw && d < tAB) { r = y + h; gv = -0.5; }
    else if (d >= tAB && d < tBC) { r = y - H / L * (d + W / 4.0); gv = (d + W / 4.0) / L; }
    else if (d >= tBC && d < tCD) { r = y - h; gv = 0.5; }
    else if (d >= tCD && d < tDE) { r = y + H / L * (d - W / 4.0); gv = -(d - W / 4.0) / L; }
    else if (d >= tDE && d < hw) { r = y + h; gv = -0.5; }
    else continue;
    res += r * r;
    ga += gv * gv;
    gb += y * gv;
  }
  const int nn = nA + nB + nC + nD + nE;
  if (Hout) *Hout = H;
  if (cnt) *cnt = nn;
  if (resout) *resout = res;
  if (meanout) *meanout = mean;
  if (nn <= 3) return NAN; /* n - 3 degrees of freedom: no trial below 4 samples */
  if (!(W < P)) return NAN; /* no trial: the window spans the period */
  if (m == 0) return NAN;   /* L = 0: no trial (tls.py's form gives 0/0) */
  double Hls = H;
  const double v = trial_value(H, res, ga, gb, nn, stat, &Hls);
  if (Hout) *Hout = Hls;
  return stat ? dchi2_value(v, flux_sigma(t, x, n)) : v;
}

/* Window-restricted version of the trial above: the same sums over the
 * samples whose offsets were precomputed (d[k], y-values via x[k]). */
static double trial_window(const double *d, const double *xv, int k, double W, int m, int nl,
                           int stat, double *Hout) {
  const double dl = 0.45 * W / (double)nl;
  const double L = (double)m * dl;
  const double hw = 0.5 * W, qw = 0.25 * W, hl = 0.5 * L;
  const double tAB = -qw - hl, tBC = -qw + hl, tCD = qw - hl, tDE = qw + hl;
  double msum = 0.0;
  int mcnt = 0;
  for (int i = 0; i < k; ++i)
    if (d[i] >= -hw && d[i] <= hw) { msum += xv[i]; mcnt++; }
  const double mean = msum / (double)mcnt;
  int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
  double xAEC = 0, xBD = 0, xtB = 0, xtD = 0, tB = 0, tD = 0, t2BD = 0;
  for (int i = 0; i < k; ++i) {
    const double dd = d[i], y = xv[i] - mean;
    if (dd >= -hw && dd < tAB) { nA++; xAEC += y; }
    else if (dd >= tAB && dd < tBC) { nB++; xBD += y; xtB += y * dd; tB += dd; t2BD += dd * dd; }
    else if (dd >= tBC && dd < tCD) { nC++; xAEC += y; }
    else if (dd >= tCD && dd < tDE) { nD++; xBD += y; xtD += y * dd; tD += dd; t2BD += dd * dd; }
    else if (dd >= tDE && dd < hw) { nE++; xAEC += y; }
  }
  const double denA = 4.0 * L * L * (double)(nA - nC + nE) + W * W * (double)(nB + nD) +
                      8.0 * W * (tB - tD) + 16.0 * t2BD;
  const double numB = 8.0 * L * L * xAEC - 4.0 * W * L * xBD + 16.0 * L * (xtD - xtB);
  const double H = -numB / denA;
  const double h = H / 2.0;
  double res = 0.0, ga = 0.0, gb = 0.0;
  for (int i = 0; i < k; ++i) {
    const double dd = d[i], y = xv[i] - mean;
    double r, gv;  /* gv: the trapezoid of unit depth, tls.py:2-15 */
    if (dd >= -hw && dd < tAB) { r = y + h; gv = -0.5; }
    else if (dd >= tAB && dd < tBC) { r = y - H / L * (dd + W / 4.0); gv = (dd + W / 4.0) / L; }
    else if (dd >= tBC && dd < tCD) { r = y - h; gv = 0.5; }
    else if (dd >= tCD && dd < tDE) { r = y + H / L * (dd - W / 4.0); gv = -(dd - W / 4.0) / L; }
    else if (dd >= tDE && dd < hw) { r = y + h; gv = -0.5; }
    else continue;
    res += r * r;
    ga += gv * gv;
    gb += y * gv;
  }
  const int nn = nA + nB + nC + nD + nE;
  if (Hout) *Hout = H;
  if (nn <= 3 || m == 0) return NAN;
  double Hls = H;
  const double v = trial_value(H, res, ga, gb, nn, stat, &Hls);
  if (Hout) *Hout = Hls;
  return v;
}

typedef struct { double u; double x; } ux_t;
static int cmp_ux(const void *a, const void *b) {
  const double ua = ((const ux_t *)a)->u, ub = ((const ux_t *)b)->u;
  return (ua > ub) - (ua < ub);
}

static int lower_u(const ux_t *s, int k, double v) { /* first index with u >= v */
  int lo = 0, hi = k;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (s[mid].u < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

void orc_tlsfold(const float *t, const float *x, int n, const double *periods, int np,
                 const double *widths, int nw, int nl, double dtep, int stat, double *best_sn,
                 int *best_w, int *best_l, int *best_e, double *best_h) {
  double wmax = 0.0;
  for (int w = 0; w < nw; ++w) if (widths[w] > wmax) wmax = widths[w];
  const double hwm = 0.5 * wmax;
  const double sig = stat ? flux_sigma(t, x, n) : 1.0;
  ux_t *s = (ux_t *)malloc(sizeof(ux_t) * (size_t)(n > 0 ? n : 1));
  double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *xv = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int p = 0; p < np; ++p) {
    const double P = periods[p];
    int k = 0;
    for (int i = 0; i < n; ++i) {
      if (!usable(t[i], x[i])) continue;
      s[k].u = fmod((double)t[i], P);
      s[k].x = (double)x[i];
      k++;
    }
    qsort(s, (size_t)k, sizeof(ux_t), cmp_ux);
    const int nE = (int)ceil(P / dtep);
    double bsn = -INFINITY, bh = NAN;
    int bw = -1, bl = -1, be = -1;
    for (int e = 0; e < nE; ++e) {
      const double t0 = (double)e * dtep;
      /* candidates: samples whose folded offset lies in [-Wmax/2, Wmax/2];
       * the sorted order only narrows the scan (1e-9 d margin), the test
       * itself uses the exact offset */
      int kk = 0;
      if (2.0 * (hwm + 1e-9) < P) {
        /* three disjoint image ranges of the sorted phases */
        for (int img = -1; img <= 1; ++img) {
          const double lo = t0 - hwm - 1e-9 - img * P, hi = t0 + hwm + 1e-9 - img * P;
          for (int i = lower_u(s, k, lo); i < k && s[i].u <= hi; ++i) {
            const double dd = fold_offset(s[i].u, t0, P);
            if (dd >= -hwm && dd <= hwm) { d[kk] = dd; xv[kk] = s[i].x; kk++; }
          }
        }
      } else {
        /* the ranges overlap (Wmax >= P): every sample once, at its offset */
        for (int i = 0; i < k; ++i) {
          const double dd = fold_offset(s[i].u, t0, P);
          if (dd >= -hwm && dd <= hwm) { d[kk] = dd; xv[kk] = s[i].x; kk++; }
        }
      }
      for (int w = 0; w < nw; ++w) {
        for (int m = 0; m < nl; ++m) {
          double H;
          const double sn = widths[w] < P ? trial_window(d, xv, kk, widths[w], m, nl, stat, &H)
                                          : (H = NAN, NAN);
          const int earlier = (w < bw) || (w == bw && (m < bl || (m == bl && e < be)));
          if (sn > bsn || (sn == bsn && earlier)) { bsn = sn; bw = w; bl = m; be = e; bh = H; }
        }
      }
    }
    best_sn[p] = (stat && bw >= 0) ? dchi2_value(bsn, sig) : bsn;
    best_w[p] = bw;
    best_l[p] = bl;
    best_e[p] = be;
    if (best_h) best_h[p] = bh;
  }
  free(s); free(d); free(xv);
}

/* ------------------------------------------------------------------------ */
/* gtls: include/gtls_simple.h:1-406 (mode 0), include/gtls_absolute.h      */
/* (mode 1; dl = W/4/nl, L = (m+1)dl, S/N uses |H|).  One "block" per epoch  */
/* k and curve q; one "thread" per width W = wmax - dw*iw.  Float32 as the   */
/* reference; expressions with double literals (FILLVAL -1000.0, 0.45, 0.5,  */
/* 0.25, 4.0, 8.0, 16.0, W/4.0, Hmax/2.0) are evaluated in double and        */
/* rounded to float where the reference assigns to a float.  a*b+c float     */
/* patterns are fused (nvcc's default -fmad=true contraction); division and  */
/* sqrt are IEEE (the reference's -use_fast_math approximations are not      */
/* reproducible bit for bit).  Ties of the block maximum resolve to the      */
/* lowest thread (largest W).                                               */
/* ------------------------------------------------------------------------ */
#define G_FILLVAL (-1000.0)
#define G_FILFACCRIT 0.75

static void gtls_thread(const float *cache, int nsc, int nl, float W, float deltat, int mode,
                        float *res_max_out, float *Lmax_out, float *Hmax_out) {
  const float *ct = cache + nsc;
  const float dl = mode ? (float)((double)W / 4.0 / (double)(float)nl)
                        : (float)((double)W * 0.45 / (double)(float)nl);   /* :69 */
  const float tEnd = (float)(0.5 * (double)W), tStart = (float)(-0.5 * (double)W); /* :116 */
  float mean = 0.0f, nsum = 0.0f;                                       /* :139-152 */
  for (int m = 0; m < nsc; ++m) {
    if (ct[m] >= tStart && ct[m] <= tEnd) {
      mean = mean + cache[m];
      nsum = (float)((double)nsum + 1.0);
    }
  }
  if (nsum > 0) mean = mean / nsum;
  float res_max = 0.0f, Lmax = 0.0f, Hmaxmax = 0.0f;
  for (int mm = 0; mm < nl; ++mm) {                                     /* :162 */
    const float L = (mode ? (float)(mm + 1) : (float)mm) * dl;
    float xAEC = 0, xBD = 0, xtB = 0, xtD = 0, t2BD = 0, tB = 0, tD = 0;
    int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
    int nSx = 0, nAx = 0, nBx = 0, nCx = 0, nDx = 0, nEx = 0;
    const float tAB = (float)((double)(-W) * 0.25 - (double)L * 0.5);  /* :191-194 */
    const float tBC = (float)((double)(-W) * 0.25 + (double)L * 0.5);
    const float tCD = -tBC, tDE = -tAB;
    for (int i = 0; i < nsc; ++i) { if (ct[i] > tStart) break; nSx++; }  /* :199-206 */
    for (int i = 0; i < nsc - nSx; ++i) {                               /* :212-225 */
      const float tnow = ct[i + nSx];
      if (tnow > tAB) break;
      if (tnow >= tStart) { nA++; xAEC = (xAEC + cache[i]) - mean; }   /* A.1 quirk */
      nAx++;
    }
    for (int i = 0; i < nsc - nAx - nSx; ++i) {                         /* :231-249 */
      const float tnow = ct[i + nSx + nAx];
      if (tnow > tBC) break;
      if (tnow >= tStart) {
        nB++;
        const float xtmp = cache[i + nSx + nAx] - mean;
        xBD = xBD + xtmp;
        tB = tB + tnow;
        xtB = fmaf(xtmp, tnow, xtB);
        t2BD = fmaf(tnow, tnow, t2BD);
      }
      nBx++;
    }
    for (int i = 0; i < nsc - nAx - nBx - nSx; ++i) {                   /* :252-268 */
      const float tnow = ct[i + nSx + nAx + nBx];
      if (tnow > tCD) break;
      if (tnow >= tStart) { nC++; xAEC = (xAEC + cache[i + nSx + nAx + nBx]) - mean; }
      nCx++;
    }
    for (int i = 0; i < nsc - nSx - nAx - nBx - nCx; ++i) {             /* :271-288 */
      const float tnow = ct[i + nSx + nAx + nBx + nCx];
      if (tnow > tDE) break;
      if (tnow >= tStart) {
        nD++;
        const float xtmp = cache[i + nSx + nAx + nBx + nCx] - mean;
        xBD = xBD + xtmp;
        tD = tD + tnow;
        xtD = fmaf(xtmp, tnow, xtD);
        t2BD = fmaf(tnow, tnow, t2BD);
      }
      nDx++;
    }
    for (int i = 0; i < nsc - nAx - nBx - nCx - nDx - nSx; ++i) {       /* :292-305 */
      const float tnow = ct[i + nSx + nAx + nBx + nCx + nDx];
      if (tnow > tEnd) break;
      if (tnow >= tStart) { nE++; xAEC = (xAEC + cache[i + nSx + nAx + nBx + nCx + nDx]) - mean; }
      nEx++;
    }
    const float filfac = (float)(nA + nB + nC + nD + nE) / (W / deltat);  /* :308 */
    if ((double)filfac > G_FILFACCRIT) {                                /* :311 */
      const double dL = (double)L, dW = (double)W;
      const float denA = (float)(4.0 * dL * dL * (double)(float)(nA - nC + nE) +
                                 (double)(W * W * (float)(nB + nD)) + 8.0 * dW * (double)(tB - tD) +
                                 16.0 * (double)t2BD);                  /* :314 */
      const float numB = (float)(8.0 * dL * dL * (double)xAEC - 4.0 * dW * dL * (double)xBD +
                                 16.0 * dL * (double)(xtD - xtB));      /* :315 */
      const float Hmax = -numB / denA;                                   /* :316 */
      const float Hmaxh = (float)((double)Hmax / 2.0);
      const float HL = Hmax / L;
      float res = 0.0f;
      for (int i = nSx; i < nSx + nAx; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) + Hmaxh; res = fmaf(e, e, res); }
      for (int i = nSx + nAx; i < nSx + nAx + nBx; ++i)
        if (ct[i] >= tStart) {
          const float e = (float)((double)(cache[i] - mean) -
                                  (double)HL * ((double)ct[i] + dW / 4.0));
          res = fmaf(e, e, res);
        }
      for (int i = nSx + nAx + nBx; i < nSx + nAx + nBx + nCx; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) - Hmaxh; res = fmaf(e, e, res); }
      for (int i = nSx + nAx + nBx + nCx; i < nSx + nAx + nBx + nCx + nDx; ++i)
        if (ct[i] >= tStart) {
          const float e = (float)((double)(cache[i] - mean) +
                                  (double)HL * ((double)ct[i] - dW / 4.0));
          res = fmaf(e, e, res);
        }
      for (int i = nSx + nAx + nBx + nCx + nDx; i < nSx + nAx + nBx + nCx + nDx + nEx; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) + Hmaxh; res = fmaf(e, e, res); }
      if (res > 0.0f) {                                                  /* :359-369 */
        const float hv = mode ? fabsf(Hmax) : Hmax;
        const float sn = hv / sqrtf(res / (float)(nA + nB + nC + nD + nE - 3));
        if (sn > res_max) { res_max = sn; Lmax = L; Hmaxmax = Hmax; }
      }
    }
  }
  *res_max_out = res_max;
  *Lmax_out = Lmax;
  *Hmax_out = Hmaxmax;
}

void orc_gtls(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
              const float *lc, const float *tu, int nt, int nq, int nw, int nl, int nsc,
              float wmax, float dw, float deltat, int mode, int k0, int k1) {
  float *cache = (float *)malloc(sizeof(float) * 2 * (size_t)(nsc > 0 ? nsc : 1));
  const int nsch = nsc / 2;                                              /* :52 */
  for (int k = k0; k < k1; ++k) {
    for (int q = 0; q < nq; ++q) {
      for (int i = 0; i < nsc; ++i) {                                    /* :28-65 */
        cache[i] = 0.0f;
        cache[i + nsc] = (float)G_FILLVAL;
        const int s = k + i - nsch;
        if (s >= 0 && s < nt) {
          cache[i] = 0.0f + lc[(size_t)s * nq + q];
          cache[i + nsc] = cache[i + nsc] + (float)((double)tu[(size_t)s * nq + q] - G_FILLVAL);
        }
      }
      for (int i = 0; i < nsc; ++i) {                                    /* :122-134 */
        if (cache[nsc + i] >= 0.0f) cache[i + nsc] = fmaf(-deltat, (float)k, cache[i + nsc]);
        else if (cache[nsc + i] < 0.0f) { cache[i + nsc] = -1000.0f; cache[i] = 0.0f; }
      }
      float best = 0.0f, bL = 0.0f, bH = 0.0f, bW = 0.0f;
      int have = 0;
      for (int iw = 0; iw < nw; ++iw) {
        const float W = fmaf(-dw, (float)iw, wmax);                       /* :68 */
        float rm, Lm, Hm;
        gtls_thread(cache, nsc, nl, W, deltat, mode, &rm, &Lm, &Hm);
        if (!have || rm > best) { best = rm; bL = Lm; bH = Hm; bW = W; have = 1; }
      }
      const size_t o = (size_t)k * nq + q;                                /* :394-404 */
      tlssn[o] = best;
      tlst0[o] = deltat * (float)k;
      tlsw[o] = bW;
      tlsl[o] = bL;
      tlshmax[o] = bH;
    }
  }
  free(cache);
}

/* ------------------------------------------------------------------------ */
/* gtls in float64 (the arithmetic of the build's fast gtls mode).  The      */
/* region membership is decided exactly as include/gtls_simple.h decides it: */
/* shifted times ct = fl32(T - deltat*k) (one rounding, a fused multiply-    */
/* add) against float32 boundaries tStart, tAB, tBC, tCD, tDE, tEnd, the    */
/* sequential scans with their breaks (:199-305), the region-A flux index   */
/* quirk (:214-221), the float32 fill factor (:308-311), the mean over the  */
/* closed window (:139-152).  Every sum, H (:314-316), the residual         */
/* (:324-357) and S/N (:359-369) are float64, with the exact shifted time   */
/* tau = T - deltat*k (float64, no rounding) in the sums.  Ties resolve to  */
/* the first L and then the lowest W index.                                 */
/* ------------------------------------------------------------------------ */
static void gtls_thread64(const float *x, const float *ct, const double *tau, int nsc, int nl,
                          float W, float deltat, int mode, double *res_max_out,
                          float *Lmax_out, double *Hmax_out) {
  const float dl = mode ? (float)((double)W / 4.0 / (double)(float)nl)
                        : (float)((double)W * 0.45 / (double)(float)nl);
  const float tEnd = (float)(0.5 * (double)W), tStart = (float)(-0.5 * (double)W);
  double msum = 0.0;
  int mcnt = 0;
  for (int m = 0; m < nsc; ++m)
    if (ct[m] >= tStart && ct[m] <= tEnd) { msum += (double)x[m]; mcnt++; }
  const double mean = mcnt > 0 ? msum / (double)mcnt : 0.0;
  int nSx = 0;
  for (int i = 0; i < nsc; ++i) { if (ct[i] > tStart) break; nSx++; }
  double res_max = 0.0, Hmaxmax = 0.0;
  float Lmax = 0.0f;
  const double dW = (double)W;
  for (int mm = 0; mm < nl; ++mm) {
    const float L = (mode ? (float)(mm + 1) : (float)mm) * dl;
    const double dL = (double)L;
    double xAEC = 0, xBD = 0, xtB = 0, xtD = 0, t2BD = 0, tB = 0, tD = 0;
    int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
    int nAx = 0, nBx = 0, nCx = 0, nDx = 0, nEx = 0;
    const float tAB = (float)((double)(-W) * 0.25 - (double)L * 0.5);
    const float tBC = (float)((double)(-W) * 0.25 + (double)L * 0.5);
    const float tCD = -tBC, tDE = -tAB;
    for (int i = 0; i < nsc - nSx; ++i) {
      if (ct[i + nSx] > tAB) break;
      if (ct[i + nSx] >= tStart) { nA++; xAEC += (double)x[i] - mean; }  /* A.1 quirk */
      nAx++;
    }
    for (int i = nSx + nAx; i < nsc; ++i) {
      if (ct[i] > tBC) break;
      if (ct[i] >= tStart) {
        const double y = (double)x[i] - mean;
        nB++; xBD += y; tB += tau[i]; xtB += y * tau[i]; t2BD += tau[i] * tau[i];
      }
      nBx++;
    }
    for (int i = nSx + nAx + nBx; i < nsc; ++i) {
      if (ct[i] > tCD) break;
      if (ct[i] >= tStart) { nC++; xAEC += (double)x[i] - mean; }
      nCx++;
    }
    for (int i = nSx + nAx + nBx + nCx; i < nsc; ++i) {
      if (ct[i] > tDE) break;
      if (ct[i] >= tStart) {
        const double y = (double)x[i] - mean;
        nD++; xBD += y; tD += tau[i]; xtD += y * tau[i]; t2BD += tau[i] * tau[i];
      }
      nDx++;
    }
    for (int i = nSx + nAx + nBx + nCx + nDx; i < nsc; ++i) {
      if (ct[i] > tEnd) break;
      if (ct[i] >= tStart) { nE++; xAEC += (double)x[i] - mean; }
      nEx++;
    }
    const int n = nA + nB + nC + nD + nE;
    const float filfac = (float)n / (W / deltat);
    if (!((double)filfac > G_FILFACCRIT)) continue;
    const double denA = 4.0 * dL * dL * (double)(nA - nC + nE) + dW * dW * (double)(nB + nD) +
                        8.0 * dW * (tB - tD) + 16.0 * t2BD;
    const double numB = 8.0 * dL * dL * xAEC - 4.0 * dW * dL * xBD + 16.0 * dL * (xtD - xtB);
    const double H = -numB / denA, Hh = H / 2.0, HL = H / dL;
    double res = 0.0;
    const int eA = nSx + nAx, eB = eA + nBx, eC = eB + nCx, eD = eC + nDx, eE = eD + nEx;
    for (int i = nSx; i < eE; ++i) {
      if (!(ct[i] >= tStart)) continue;
      const double y = (double)x[i] - mean;
      double e;
      if (i < eA) e = y + Hh;
      else if (i < eB) e = y - HL * (tau[i] + dW / 4.0);
      else if (i < eC) e = y - Hh;
      else if (i < eD) e = y + HL * (tau[i] - dW / 4.0);
      else e = y + Hh;
      res += e * e;
    }
    if (res > 0.0) {
      const double hv = mode ? fabs(H) : H;
      const double sn = hv / sqrt(res / (double)(n - 3));
      if (sn > res_max) { res_max = sn; Lmax = L; Hmaxmax = H; }
    }
  }
  *res_max_out = res_max;
  *Lmax_out = Lmax;
  *Hmax_out = Hmaxmax;
}

void orc_gtls64(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
                double *sn64, const float *lc, const float *tu, int nt, int nq, int nw, int nl,
                int nsc, float wmax, float dw, float deltat, int mode, int k0, int k1) {
  float *x = (float *)malloc(sizeof(float) * (size_t)(nsc > 0 ? nsc : 1));
  float *ct = (float *)malloc(sizeof(float) * (size_t)(nsc > 0 ? nsc : 1));
  double *tau = (double *)malloc(sizeof(double) * (size_t)(nsc > 0 ? nsc : 1));
  const int nsch = nsc / 2;
  for (int k = k0; k < k1; ++k) {
    for (int q = 0; q < nq; ++q) {
      for (int i = 0; i < nsc; ++i) {
        float T = (float)G_FILLVAL, f = 0.0f;
        const int s = k + i - nsch;
        if (s >= 0 && s < nt) {
          f = lc[(size_t)s * nq + q];
          T = T + (float)((double)tu[(size_t)s * nq + q] - G_FILLVAL);
        }
        if (T >= 0.0f) {
          ct[i] = fmaf(-deltat, (float)k, T);
          tau[i] = (double)T - (double)deltat * (double)k;
          x[i] = f;
        } else {
          ct[i] = -1000.0f;
          tau[i] = 0.0;
          x[i] = 0.0f;
        }
      }
      double best = 0.0, bH = 0.0;
      float bL = 0.0f, bW = 0.0f;
      int have = 0;
      for (int iw = 0; iw < nw; ++iw) {
        const float W = fmaf(-dw, (float)iw, wmax);
        double rm, Hm;
        float Lm;
        gtls_thread64(x, ct, tau, nsc, nl, W, deltat, mode, &rm, &Lm, &Hm);
        if (!have || rm > best) { best = rm; bL = Lm; bH = Hm; bW = W; have = 1; }
      }
      const size_t o = (size_t)k * nq + q;
      tlssn[o] = (float)best;
      if (sn64) sn64[o] = best;
      tlst0[o] = deltat * (float)k;
      tlsw[o] = bW;
      tlsl[o] = bL;
      tlshmax[o] = (float)bH;
    }
  }
  free(x); free(ct); free(tau);
}
