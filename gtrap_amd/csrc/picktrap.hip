// Candidate cut-outs: the W-normalised, gap-cleaned, resampled light-curve
// vector around a candidate transit (gtrap/picktrap.py:27-139,
// pick_Wnormalized_cleaned_lc_direct), for a batch of candidates (curve q,
// T0, W) on the interleaved (n, nq) image.
//
// Per candidate: ii = the first slot with tu > 0 and tu >= T0 (float64
// comparison); h = int(1.11 alpha int(alpha W daytopix)); window slots
// [ii-h, ii+h), outside the curve time 1000 and the window's NaN-free
// median flux (float32 arithmetic); prec as :66-77; Jacobi gap-cleaning
// sweeps (:82-112: a gap between two valid (> 0) slots takes their mean, next
// to one valid slot the median of the window's current fluxes, then time 1000
// and igap 1) until no gap is left; flux normalised by its mean and
// population std; resampled onto tx = i*2 alpha/nx - alpha from
// tt = k*2 alphad/nlcs - alphad (alphad = 1.5 alpha) with numpy's interp
// formula for the flux and interp1d's searchsorted formula for igap; a tx
// outside [tt[0], tt[-1]] gives all ones.
//
// status: 0 ok, 1 no valid time >= T0, 2 cleaning stuck (a gap without a
// valid neighbour; the reference loops for ever), 3 window shorter than 2,
// 4 window longer than kCtMax slots, 5 curve index outside [0, nq).
//
// MI355X design: one 256-thread workgroup per candidate (persistent over
// candidates); the window lives in LDS; medians by a bitonic sort in LDS;
// means and standard deviations by float64 block reductions.
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kCtThreads = 256;
constexpr int kCtMax = 2048;

struct CtArgs {
  const float *lc, *tu;
  int32_t n, nq;
  const int32_t *cq;
  const double *ct0, *cw;
  int32_t ncand;
  double alpha, daytopix;
  int32_t nx, contcrit;
  double *lcsx, *igapx;
  int32_t *prec, *status;
};

__device__ void ct_sort(double *a, int n) {  // ascending, n a power of two
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += kCtThreads) {
        const int l = i ^ j;
        if (l > i) {
          const double x = a[i], y = a[l];
          if ((x > y) == ((i & k) == 0)) { a[i] = y; a[l] = x; }
        }
      }
      __syncthreads();
    }
}

__device__ double ct_sum(double v, double *red) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kCtThreads / kWave; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ double ct_tt(int k, double alphad, int nl) {
  return (double)k * 2.0 * alphad / (double)nl - alphad;
}

__global__ __launch_bounds__(kCtThreads) void cutout_kernel(CtArgs A) {
  __shared__ double lcs[kCtMax], lcs2[kCtMax], srt[kCtMax];
  __shared__ float tus[kCtMax], tus2[kCtMax];
  __shared__ int igap[kCtMax];
  __shared__ double red[kCtThreads / kWave];
  __shared__ int s_min, s_prec, s_status;
  const int tid = threadIdx.x, n = A.n, nq = A.nq, nx = A.nx;
  for (int c = blockIdx.x; c < A.ncand; c += gridDim.x) {
    const int qin = A.cq[c];
    const bool qok = qin >= 0 && qin < nq;
    const int q = qok ? qin : 0;
    const double T0 = A.ct0[c], W = A.cw[c];
    if (tid == 0) { s_min = n; s_status = 0; s_prec = 1; }
    __syncthreads();
    int mk = n;
    for (int k = tid; k < n; k += kCtThreads) {
      const float t = A.tu[(size_t)k * nq + q];
      if (t > 0.0f && (double)t >= T0) { mk = k; break; }
    }
    atomicMin(&s_min, mk);
    __syncthreads();
    const int ii = s_min;
    const int wid = (int)(A.alpha * W * A.daytopix);
    const int h = (int)(1.11 * A.alpha * (double)wid);
    const int nget = 2 * h;
    int status = ii >= n ? 1 : (nget < 2 ? 3 : (nget > kCtMax ? 4 : 0));
    if (!qok) status = 5;
    double *ox = A.lcsx + (size_t)c * nx, *og = A.igapx + (size_t)c * nx;
    if (status) {
      for (int i = tid; i < nx; i += kCtThreads) { ox[i] = 1.0; og[i] = 1.0; }
      if (tid == 0) { A.prec[c] = 0; A.status[c] = status; }
      __syncthreads();
      continue;
    }
    const int iss = ii - h;
    const int istart = iss < 0 ? 0 : iss, iend = min(ii + h, n);
    // window median of the curve's slots (NaN excluded), float32 arithmetic
    const int m0 = iend - istart;
    int npad = 1;
    while (npad < m0) npad <<= 1;
    int nnan = 0;
    for (int k = tid; k < npad; k += kCtThreads) {
      double v = INFINITY;
      if (k < m0) {
        const float f = A.lc[(size_t)(istart + k) * nq + q];
        if (f == f) v = (double)f; else ++nnan;
      }
      srt[k] = v;
    }
    nnan = (int)ct_sum((double)nnan, red);
    ct_sort(srt, npad);
    const int m = m0 - nnan;
    float med = __int_as_float(0x7fc00000);
    if (m > 0) {
      if (m & 1) med = (float)srt[m / 2];
      else { const float s = (float)srt[m / 2 - 1] + (float)srt[m / 2]; med = s / 2.0f; }
    }
    __syncthreads();
    for (int k = tid; k < nget; k += kCtThreads) {
      const int src = iss + k;
      const bool in = src >= 0 && src < n;
      tus[k] = in ? A.tu[(size_t)src * nq + q] : 1000.0f;
      lcs[k] = in ? (double)A.lc[(size_t)src * nq + q] : (double)med;
      igap[k] = 0;
    }
    __syncthreads();
    if (tid == 0 && (tus[0] < 0.0f || tus[nget - 1] < -1.0f)) {
      int run = 0;
      for (int k = 0; k < nget; ++k) {
        run = tus[k] < 0.0f ? run + 1 : 0;
        if (run > A.contcrit) s_prec = 0;
      }
    }
    // Jacobi cleaning sweeps
    int npg = 1;
    while (npg < nget) npg <<= 1;
    for (;;) {
      int ngap = 0, none = 0, nnanl = 0;
      for (int j = tid; j < nget; j += kCtThreads) {
        if (tus[j] < 0.0f) {
          ++ngap;
          const float tp = j > 0 ? tus[j - 1] : -2000.0f, tm = j < nget - 1 ? tus[j + 1] : -2000.0f;
          if (!(tm > 0.0f && tp > 0.0f) && (tm > 0.0f || tp > 0.0f)) ++none;
        }
        if (!(lcs[j] == lcs[j])) ++nnanl;
      }
      ngap = (int)ct_sum((double)ngap, red);
      if (ngap == 0) break;
      none = (int)ct_sum((double)none, red);
      double medv = 0.0;
      if (none) {  // median of the current window (NaN if any flux is NaN)
        nnanl = (int)ct_sum((double)nnanl, red);
        for (int k = tid; k < npg; k += kCtThreads) srt[k] = k < nget ? lcs[k] : INFINITY;
        __syncthreads();
        ct_sort(srt, npg);
        medv = nnanl ? __longlong_as_double(0x7ff8000000000000LL)
                     : ((nget & 1) ? srt[nget / 2] : (srt[nget / 2 - 1] + srt[nget / 2]) / 2.0);
      }
      int filled = 0;
      for (int j = tid; j < nget; j += kCtThreads) {
        float t = tus[j];
        double v = lcs[j];
        if (t < 0.0f) {
          const float tp = j > 0 ? tus[j - 1] : -2000.0f, tm = j < nget - 1 ? tus[j + 1] : -2000.0f;
          if (tm > 0.0f && tp > 0.0f) {
            v = (lcs[j + 1] + lcs[j - 1]) / 2.0;
          } else if (tm > 0.0f || tp > 0.0f) {
            v = medv;
          }
          if (tm > 0.0f || tp > 0.0f) { t = 1000.0f; igap[j] = 1; ++filled; }
        }
        tus2[j] = t;
        lcs2[j] = v;
      }
      filled = (int)ct_sum((double)filled, red);
      for (int j = tid; j < nget; j += kCtThreads) { tus[j] = tus2[j]; lcs[j] = lcs2[j]; }
      __syncthreads();
      if (filled == 0) { status = 2; break; }
    }
    // normalise by the mean and population std
    double s = 0.0;
    for (int j = tid; j < nget; j += kCtThreads) s += lcs[j];
    const double mean = ct_sum(s, red) / nget;
    double s2 = 0.0;
    for (int j = tid; j < nget; j += kCtThreads) { const double d = lcs[j] - mean; s2 += d * d; }
    const double sd = sqrt(ct_sum(s2, red) / nget);
    for (int j = tid; j < nget; j += kCtThreads) lcs[j] = (lcs[j] - mean) / sd;
    __syncthreads();
    // resample
    const double alphad = A.alpha * 1.5;
    const int nl = nget;
    const double tt0 = ct_tt(0, alphad, nl), ttl = ct_tt(nl - 1, alphad, nl);
    const double tx0 = -A.alpha, txl = (double)(nx - 1) * 2.0 * A.alpha / (double)nx - A.alpha;
    const bool inb = !(tx0 < tt0 || txl > ttl);
    for (int i = tid; i < nx; i += kCtThreads) {
      if (!inb) { ox[i] = 1.0; og[i] = 1.0; continue; }
      const double x = (double)i * 2.0 * A.alpha / (double)nx - A.alpha;
      // np.interp: the segment [tt[j], tt[j+1]) holding x
      int j = (int)floor((x + alphad) * (double)nl / (2.0 * alphad));
      j = j < 0 ? 0 : (j > nl - 1 ? nl - 1 : j);
      while (j + 1 < nl && ct_tt(j + 1, alphad, nl) <= x) ++j;
      while (j > 0 && ct_tt(j, alphad, nl) > x) --j;
      double y;
      if (x == ttl) {
        y = lcs[nl - 1];
      } else {
        const int jj = j > nl - 2 ? nl - 2 : j;
        const double xl = ct_tt(jj, alphad, nl), xh = ct_tt(jj + 1, alphad, nl);
        const double slope = (lcs[jj + 1] - lcs[jj]) / (xh - xl);
        y = slope * (x - xl) + lcs[jj];
        if (!(y == y)) y = slope * (x - xh) + lcs[jj + 1];
      }
      ox[i] = y;
      // interp1d on integers: searchsorted(tt, x, 'left') clipped to [1, nl-1]
      int k = ct_tt(j, alphad, nl) < x ? j + 1 : j;
      k = k < 1 ? 1 : (k > nl - 1 ? nl - 1 : k);
      const double gl = (double)igap[k - 1], gh = (double)igap[k];
      const double xl = ct_tt(k - 1, alphad, nl), xh = ct_tt(k, alphad, nl);
      og[i] = (gh - gl) / (xh - xl) * (x - xl) + gl;
    }
    if (tid == 0) { A.prec[c] = s_prec; A.status[c] = status; }
    __syncthreads();
  }
}

}  // namespace gtrap

extern "C" int gtrap_cutout(double *lcsx, double *igapx, int32_t *prec, int32_t *status,
                            const float *lc, const float *tu, int32_t n, int32_t nq,
                            const int32_t *cq, const double *ct0, const double *cw,
                            int32_t ncand, double alpha, int32_t nx, double daytopix,
                            int32_t contcrit, void *stream) {
  using namespace gtrap;
  GTRAP_REQUIRE(lcsx && igapx && prec && status && lc && tu && cq && ct0 && cw,
                "gtrap_cutout: null pointer");
  GTRAP_REQUIRE(n >= 1 && nq >= 1 && ncand >= 0 && nx >= 1 && alpha > 0.0 && daytopix > 0.0,
                "gtrap_cutout: bad sizes or parameters");
  if (ncand == 0) return GTRAP_OK;
  CtArgs A{lc, tu, n, nq, cq, ct0, cw, ncand, alpha, daytopix, nx, contcrit, lcsx, igapx, prec,
           status};
  const int g = ncand < 2048 ? ncand : 2048;
  hipLaunchKernelGGL(cutout_kernel, dim3((unsigned)g), dim3(kCtThreads), 0, as_stream(stream), A);
  return check_launch("cutout_kernel");
}
