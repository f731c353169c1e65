// Trapezoid least squares over an explicit (W, L, t0) grid: gtrap/tls.py's
// tlsfit (tls.py:82-110) on the GPU.
//
// One wave per trial (w, l, k), k fastest.  Each trial is evaluated exactly
// as get_chi2Hmax (tls.py:44-79): tn = t - t0; regions A [-W/2, -W/4-L/2),
// B [-W/4-L/2, -W/4+L/2), C [-W/4+L/2, W/4-L/2), D [W/4-L/2, W/4+L/2),
// E [W/4+L/2, W/2) (independent masks: for L > W/2 B and D overlap); with
// divx the mean over [-W/2, W/2] is removed first;
// denA, numB, H = -numB/denA, res = sum of squared residuals / sigma^2.
// float64 throughout, direct sums over the samples (NaN samples skipped, as
// numpy's nansum/nanmean).  The wave's lanes stride over the samples and
// their partial sums are reduced across the wave by a DPP scan, every lane
// taking lane 63's total (the 17 reductions through LDS shuffles took the
// 1e6-trial scan 2.95 ms, the DPP scan 1.88).  When the caller states that t is non-decreasing, only the samples
// with min(-W/2, tAB) <= t - t0 <= max(W/2, tDE) are visited (for L > W/2 the
// ingress region starts before -W/2 and the egress region ends after W/2, as
// tls.py's masks have it): two binary searches with those same float64 tests
// (monotone in t) find them, so every sample a region or the mean counts is
// visited.
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kTgThreads = 256;
constexpr int kTgWaves = kTgThreads / kWave;

// lane move of a double by a DPP control; lanes without a source read +0.0
template <int CTRL, int ROWS>
__device__ __forceinline__ double tg_dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// sum over the wave: an inclusive scan with DPP moves (row_shr 1, 2, 4, 8,
// then row_bcast 15 and 31), every lane taking lane 63's total
__device__ __forceinline__ double tg_sum(double v) {
  v += tg_dpp<0x111, 0xf>(v);
  v += tg_dpp<0x112, 0xf>(v);
  v += tg_dpp<0x114, 0xf>(v);
  v += tg_dpp<0x118, 0xf>(v);
  v += tg_dpp<0x142, 0xa>(v);
  v += tg_dpp<0x143, 0xc>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// First index in [a, b) with t[i] - t0 >= v (strict: > v), b if none, for
// non-decreasing t: a 64-ary search, each lane testing one pivot per round
// (pivots at or past b count as passing), so about log64(b - a) dependent
// loads instead of log2(b - a).  Wave-uniform.
__device__ __forceinline__ int tg_lower_bound(const double *__restrict__ t, int a, int b,
                                              double t0, double v, bool strict) {
  const int lane = threadIdx.x & (kWave - 1);
  while (a < b) {
    const int step = (b - a + kWave - 1) / kWave;
    const int i = a + lane * step;
    bool p = true;
    if (i < b) {
      const double d = t[i] - t0;
      p = strict ? d > v : d >= v;
    }
    const unsigned long long m = __ballot(p);
    if (m == 0ull) {  // every pivot fails: the answer lies past the last one
      a = a + (kWave - 1) * step + 1;
      continue;
    }
    const int f = __ffsll((long long)m) - 1;
    if (f == 0) return a;
    const int na = a + (f - 1) * step + 1, nb = a + f * step;
    a = na;
    b = nb < b ? nb : b;
  }
  return a;
}

__global__ __launch_bounds__(kTgThreads) void tlsgrid_kernel(
    const double *__restrict__ t, const double *__restrict__ x, int n,
    const double *__restrict__ Wl, int nW, const double *__restrict__ Ll, int nL,
    const double *__restrict__ tl, int nT, int divx, double sigma, int sorted,
    double *__restrict__ H_out, double *__restrict__ res_out, int32_t *__restrict__ n_out,
    double *__restrict__ xsum_out) {
  const int lane = threadIdx.x & (kWave - 1);
  const long trial = (long)blockIdx.x * kTgWaves + threadIdx.x / kWave;
  const long ntr = (long)nW * nL * nT;
  if (trial >= ntr) return;  // the whole wave: the kernel has no barrier
  const int k = (int)(trial % nT);
  const int l = (int)((trial / nT) % nL);
  const int w = (int)(trial / ((long)nT * nL));
  const double W = Wl[w], L = Ll[l], t0 = tl[k];
  const double hw = W / 2.0, tAB = -W / 4.0 - L / 2.0, tBC = -W / 4.0 + L / 2.0,
               tCD = W / 4.0 - L / 2.0, tDE = W / 4.0 + L / 2.0;
  int lo = 0, hi = n;
  if (sorted) {
    const double vlo = fmin(-hw, tAB), vhi = fmax(hw, tDE);
    lo = tg_lower_bound(t, 0, n, t0, vlo, false);
    hi = tg_lower_bound(t, lo, n, t0, vhi, true);
  }
  // pass 1: mean over the closed window (nanmean)
  double msum = 0.0, mcnt = 0.0;
  if (divx)
    for (int i = lo + lane; i < hi; i += kWave) {
      const double tn = t[i] - t0, xi = x[i];
      if (tn >= -hw && tn <= hw && xi == xi) { msum += xi; mcnt += 1.0; }
    }
  msum = tg_sum(msum);
  mcnt = tg_sum(mcnt);
  const double mean = divx ? msum / mcnt : 0.0;
  // pass 2: region sums
  double nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
  double sA = 0, sB = 0, sC = 0, sD = 0, sE = 0, xtB = 0, xtD = 0, tB = 0, tD = 0, t2 = 0;
  for (int i = lo + lane; i < hi; i += kWave) {
    const double tn = t[i] - t0;
    const double y = x[i] - mean;
    const bool ok = y == y;  // nansum skips NaN terms; counts use every sample
    // independent masks, as tls.py's: for L > W/2 regions B and D overlap and
    // a sample in both counts in both
    if (tn >= -hw && tn < tAB) { nA += 1; if (ok) sA += y; }
    if (tn >= tAB && tn < tBC) {
      nB += 1; tB += tn; t2 += tn * tn;
      if (ok) { sB += y; xtB += y * tn; }
    }
    if (tn >= tBC && tn < tCD) { nC += 1; if (ok) sC += y; }
    if (tn >= tCD && tn < tDE) {
      nD += 1; tD += tn; t2 += tn * tn;
      if (ok) { sD += y; xtD += y * tn; }
    }
    if (tn >= tDE && tn < hw) { nE += 1; if (ok) sE += y; }
  }
  nA = tg_sum(nA); nB = tg_sum(nB); nC = tg_sum(nC); nD = tg_sum(nD); nE = tg_sum(nE);
  sA = tg_sum(sA); sB = tg_sum(sB); sC = tg_sum(sC); sD = tg_sum(sD); sE = tg_sum(sE);
  xtB = tg_sum(xtB); xtD = tg_sum(xtD); tB = tg_sum(tB); tD = tg_sum(tD); t2 = tg_sum(t2);
  const double xAEC = sA + sE + sC, xBD = sB + sD;
  const double denA = 4 * L * L * (nA - nC + nE) + W * W * (nB + nD) + 8 * W * (tB - tD) + 16 * t2;
  const double numB = 8 * L * L * xAEC - 4 * W * L * xBD + 16 * L * (xtD - xtB);
  const double H = -numB / denA;
  const double Hh = H / 2.0;
  // pass 3: residual
  double res = 0.0;
  for (int i = lo + lane; i < hi; i += kWave) {
    const double tn = t[i] - t0;
    const double y = x[i] - mean;
    double r;
    if (tn >= -hw && tn < tAB) { r = y + Hh; if (r == r) res += r * r; }
    if (tn >= tAB && tn < tBC) { r = y - H / L * (tn + W / 4.0); if (r == r) res += r * r; }
    if (tn >= tBC && tn < tCD) { r = y - Hh; if (r == r) res += r * r; }
    if (tn >= tCD && tn < tDE) { r = y + H / L * (tn - W / 4.0); if (r == r) res += r * r; }
    if (tn >= tDE && tn < hw) { r = y + Hh; if (r == r) res += r * r; }
  }
  res = tg_sum(res);
  if (lane == 0) {
    H_out[trial] = H;
    res_out[trial] = res / (sigma * sigma);
    n_out[trial] = (int32_t)(nA + nB + nC + nD + nE);
    if (xsum_out) xsum_out[trial] = divx ? mean : 0.0;
  }
}

// tlsfit's choice over the grid (tls.py:100-110): sn = -H / sqrt(res/(n-3))
// per trial in float64 as numpy evaluates it (n - 3 in int32, then the true
// division), and numpy's nanargmax: NaN counts as -inf, the first maximum
// wins, and a grid of NaN only has no choice (index -1).  One workgroup;
// each thread keeps the first maximum of its trials (in index order), then
// the threads' maxima meet by value, ties to the lower index.
constexpr int kTbThreads = 1024;
__global__ __launch_bounds__(kTbThreads) void tlsfit_best_kernel(
    const double *__restrict__ H, const double *__restrict__ res, const int32_t *__restrict__ nn,
    const double *__restrict__ xs, int32_t ntr, double *__restrict__ sn_out,
    double *__restrict__ best) {
  __shared__ double s_v[kTbThreads / kWave];
  __shared__ int s_i[kTbThreads / kWave];
  __shared__ int s_any[kTbThreads / kWave];
  double bv = -INFINITY;
  int bi = -1, any = 0;
  for (int i = threadIdx.x; i < ntr; i += kTbThreads) {
    const double s = -H[i] / sqrt(res[i] / (double)(nn[i] - 3));
    if (sn_out) sn_out[i] = s;
    any |= isnan(s) ? 0 : 1;
    const double v = isnan(s) ? -INFINITY : s;
    if (bi < 0 || v > bv) { bv = v; bi = i; }
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const double ov = __shfl_xor(bv, o);
    const int oi = __shfl_xor(bi, o);
    any |= __shfl_xor(any, o);
    if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) { bv = ov; bi = oi; }
  }
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  if (lane == 0) { s_v[wv] = bv; s_i[wv] = bi; s_any[wv] = any; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTbThreads / kWave; ++w) {
      const double ov = s_v[w];
      const int oi = s_i[w];
      any |= s_any[w];
      if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) { bv = ov; bi = oi; }
    }
    if (!any) bi = -1;
    best[0] = (double)bi;
    best[1] = bi >= 0 ? -H[bi] / sqrt(res[bi] / (double)(nn[bi] - 3)) : NAN;
    best[2] = bi >= 0 ? H[bi] : NAN;
    best[3] = bi >= 0 && xs ? xs[bi] : 0.0;
  }
}

}  // namespace gtrap

extern "C" int gtrap_tlsfit_best(const double *H, const double *res, const int32_t *nout,
                                 const double *xsum, int32_t ntr, double *sn, double *best,
                                 void *stream) {
  GTRAP_REQUIRE(H && res && nout && best, "gtrap_tlsfit_best: null pointer");
  GTRAP_REQUIRE(ntr >= 0, "gtrap_tlsfit_best: negative size");
  hipLaunchKernelGGL(gtrap::tlsfit_best_kernel, dim3(1), dim3(gtrap::kTbThreads), 0,
                     gtrap::as_stream(stream), H, res, nout, xsum, ntr, sn, best);
  return gtrap::check_launch("tlsfit_best_kernel");
}

extern "C" int gtrap_tlsgrid(const double *t, const double *x, int32_t n, const double *Wl,
                             int32_t nW, const double *Ll, int32_t nL, const double *tl,
                             int32_t nT, int divx, double sigma, int32_t sorted, double *H,
                             double *res, int32_t *nout, double *xsum, void *stream) {
  GTRAP_REQUIRE(t && x && Wl && Ll && tl && H && res && nout, "gtrap_tlsgrid: null pointer");
  GTRAP_REQUIRE(n >= 0 && nW >= 0 && nL >= 0 && nT >= 0, "gtrap_tlsgrid: negative size");
  const long ntr = (long)nW * nL * nT;
  if (ntr == 0) return GTRAP_OK;
  GTRAP_REQUIRE(ntr < (1L << 31) * (long)gtrap::kTgWaves, "gtrap_tlsgrid: grid too large");
  hipLaunchKernelGGL(gtrap::tlsgrid_kernel,
                     dim3((unsigned)((ntr + gtrap::kTgWaves - 1) / gtrap::kTgWaves)),
                     dim3(gtrap::kTgThreads), 0, gtrap::as_stream(stream), t, x, n, Wl, nW, Ll,
                     nL, tl, nT, divx, sigma, sorted ? 1 : 0, H, res, nout, xsum);
  return gtrap::check_launch("tlsgrid_kernel");
}
