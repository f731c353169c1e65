// Trapezoid least squares over an explicit (W, L, t0) grid: gtrap/tls.py's
// tlsfit (tls.py:82-110) on the GPU.
//
// One thread per trial (w, l, k), k fastest.  Each trial is evaluated exactly
// as get_chi2Hmax (tls.py:44-79): tn = t - t0; regions A [-W/2, -W/4-L/2),
// B [-W/4-L/2, -W/4+L/2), C [-W/4+L/2, W/4-L/2), D [W/4-L/2, W/4+L/2),
// E [W/4+L/2, W/2); with divx the mean over [-W/2, W/2] is removed first;
// denA, numB, H = -numB/denA, res = sum of squared residuals / sigma^2.
// float64 throughout, direct sums over the samples (NaN samples skipped, as
// numpy's nansum/nanmean).  The curve is staged in LDS in tiles.
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kTgThreads = 256;
constexpr int kTgTile = 1024;

__global__ __launch_bounds__(kTgThreads) void tlsgrid_kernel(
    const double *__restrict__ t, const double *__restrict__ x, int n,
    const double *__restrict__ Wl, int nW, const double *__restrict__ Ll, int nL,
    const double *__restrict__ tl, int nT, int divx, double sigma, double *__restrict__ H_out,
    double *__restrict__ res_out, int32_t *__restrict__ n_out, double *__restrict__ xsum_out) {
  __shared__ double st[kTgTile], sx[kTgTile];
  const long trial = (long)blockIdx.x * kTgThreads + threadIdx.x;
  const long ntr = (long)nW * nL * nT;
  const bool live = trial < ntr;
  const int k = live ? (int)(trial % nT) : 0;
  const int l = live ? (int)((trial / nT) % nL) : 0;
  const int w = live ? (int)(trial / ((long)nT * nL)) : 0;
  const double W = Wl[w], L = Ll[l], t0 = tl[k];
  const double hw = W / 2.0, tAB = -W / 4.0 - L / 2.0, tBC = -W / 4.0 + L / 2.0,
               tCD = W / 4.0 - L / 2.0, tDE = W / 4.0 + L / 2.0;
  // pass 1: mean over the closed window (nanmean)
  double msum = 0.0, mcnt = 0.0;
  for (int a = 0; a < n; a += kTgTile) {
    const int m = min(kTgTile, n - a);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kTgThreads) { st[i] = t[a + i]; sx[i] = x[a + i]; }
    __syncthreads();
    if (live && divx)
      for (int i = 0; i < m; ++i) {
        const double tn = st[i] - t0;
        if (tn >= -hw && tn <= hw && sx[i] == sx[i]) { msum += sx[i]; mcnt += 1.0; }
      }
  }
  const double mean = divx ? msum / mcnt : 0.0;
  // pass 2: region sums
  double nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
  double sA = 0, sB = 0, sC = 0, sD = 0, sE = 0, xtB = 0, xtD = 0, tB = 0, tD = 0, t2 = 0;
  for (int a = 0; a < n; a += kTgTile) {
    const int m = min(kTgTile, n - a);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kTgThreads) { st[i] = t[a + i]; sx[i] = x[a + i]; }
    __syncthreads();
    if (live)
      for (int i = 0; i < m; ++i) {
        const double tn = st[i] - t0;
        const double y = sx[i] - mean;
        const bool ok = y == y;  // nansum skips NaN terms; counts use every sample
        if (tn >= -hw && tn < tAB) { nA += 1; if (ok) sA += y; }
        else if (tn >= tAB && tn < tBC) {
          nB += 1; tB += tn; t2 += tn * tn;
          if (ok) { sB += y; xtB += y * tn; }
        } else if (tn >= tBC && tn < tCD) { nC += 1; if (ok) sC += y; }
        else if (tn >= tCD && tn < tDE) {
          nD += 1; tD += tn; t2 += tn * tn;
          if (ok) { sD += y; xtD += y * tn; }
        } else if (tn >= tDE && tn < hw) { nE += 1; if (ok) sE += y; }
      }
  }
  const double xAEC = sA + sE + sC, xBD = sB + sD;
  const double denA = 4 * L * L * (nA - nC + nE) + W * W * (nB + nD) + 8 * W * (tB - tD) + 16 * t2;
  const double numB = 8 * L * L * xAEC - 4 * W * L * xBD + 16 * L * (xtD - xtB);
  const double H = -numB / denA;
  const double Hh = H / 2.0;
  // pass 3: residual
  double res = 0.0;
  for (int a = 0; a < n; a += kTgTile) {
    const int m = min(kTgTile, n - a);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kTgThreads) { st[i] = t[a + i]; sx[i] = x[a + i]; }
    __syncthreads();
    if (live)
      for (int i = 0; i < m; ++i) {
        const double tn = st[i] - t0;
        const double y = sx[i] - mean;
        double r;
        if (tn >= -hw && tn < tAB) r = y + Hh;
        else if (tn >= tAB && tn < tBC) r = y - H / L * (tn + W / 4.0);
        else if (tn >= tBC && tn < tCD) r = y - Hh;
        else if (tn >= tCD && tn < tDE) r = y + H / L * (tn - W / 4.0);
        else if (tn >= tDE && tn < hw) r = y + Hh;
        else continue;
        if (r == r) res += r * r;
      }
  }
  if (live) {
    H_out[trial] = H;
    res_out[trial] = res / (sigma * sigma);
    n_out[trial] = (int32_t)(nA + nB + nC + nD + nE);
    if (xsum_out) xsum_out[trial] = divx ? mean : 0.0;
  }
}

}  // namespace gtrap

extern "C" int gtrap_tlsgrid(const double *t, const double *x, int32_t n, const double *Wl,
                             int32_t nW, const double *Ll, int32_t nL, const double *tl,
                             int32_t nT, int divx, double sigma, double *H, double *res,
                             int32_t *nout, double *xsum, void *stream) {
  GTRAP_REQUIRE(t && x && Wl && Ll && tl && H && res && nout, "gtrap_tlsgrid: null pointer");
  GTRAP_REQUIRE(n >= 0 && nW >= 0 && nL >= 0 && nT >= 0, "gtrap_tlsgrid: negative size");
  const long ntr = (long)nW * nL * nT;
  if (ntr == 0) return GTRAP_OK;
  GTRAP_REQUIRE(ntr < (1L << 31) * 256L, "gtrap_tlsgrid: grid too large");
  hipLaunchKernelGGL(gtrap::tlsgrid_kernel,
                     dim3((unsigned)((ntr + gtrap::kTgThreads - 1) / gtrap::kTgThreads)),
                     dim3(gtrap::kTgThreads), 0, gtrap::as_stream(stream), t, x, n, Wl, nW, Ll,
                     nL, tl, nT, divx, sigma, H, res, nout, xsum);
  return gtrap::check_launch("tlsgrid_kernel");
}
