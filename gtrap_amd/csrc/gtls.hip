// Single-transit-epoch trapezoid least squares (gtls) for gfx950.
//
// Reference semantics: include/gtls_simple.h:1-406 (mode 0) and
// include/gtls_absolute.h (mode 1), module factory gtrap/gtls.py:11-27
// (FILLVAL -1000, FILFACCRIT 0.75).  For epoch k and curve q a scoop of nsc
// samples centred on k is staged (time stored as -1000 + fl(tu + 1000),
// gtls_simple.h:32,62), shifted by deltat*k, masked samples set to time
// -1000 and flux 0 (:121-134); every width W = wmax - dw*iw scans the scoop
// sequentially into regions A..E for each ingress L (:162-305), keeps trials
// with fill factor > 0.75, forms H = -numB/denA (:314-316) and the residual
// (:324-357) and S/N = H/sqrt(res/(n-3)) (|H| in absolute mode); the best
// over (W, L) is written with its W, L, H and t0 = deltat*k (:383-404).
//
// The float32 operation order of the reference is kept (including region A
// summing the flux at scoop index i while testing the time at i+nSx,
// :214-221), with a*b+c patterns fused and IEEE division/square root, so the
// result is bit-identical to the CPU restatement.  MI355X layout: one
// workgroup of 256 threads covers several consecutive epochs of one curve
// (thread = (epoch, W)), each epoch's scoop lives in LDS, and the block
// maximum uses a deterministic lowest-W-index tie-break (the reference lets
// every tied thread write, :398-404).
#include "common.h"
#include <math.h>
#include <type_traits>

namespace gtrap {

constexpr int kGtThreads = 256;
constexpr double kFill = -1000.0;
constexpr double kFilfac = 0.75;

__device__ void gtls_thread(const float *cache, int nsc, int nl, float W, float deltat,
                            int mode, float &res_max, float &Lmax, float &Hmaxmax) {
  const float *ct = cache + nsc;
  const float dl = mode ? (float)((double)W / 4.0 / (double)(float)nl)
                        : (float)((double)W * 0.45 / (double)(float)nl);
  const float tEnd = (float)(0.5 * (double)W), tStart = (float)(-0.5 * (double)W);
  float mean = 0.0f, nsum = 0.0f;
  for (int m = 0; m < nsc; ++m) {
    if (ct[m] >= tStart && ct[m] <= tEnd) {
      mean = mean + cache[m];
      nsum = (float)((double)nsum + 1.0);
    }
  }
  if (nsum > 0) mean = mean / nsum;
  // samples before -W/2 (the reference recounts them for every L; the count
  // does not depend on L)
  int nSx = 0;
  for (int i = 0; i < nsc; ++i) { if (ct[i] > tStart) break; nSx++; }
  res_max = 0.0f;
  Lmax = 0.0f;
  Hmaxmax = 0.0f;
  for (int mm = 0; mm < nl; ++mm) {
    const float L = (mode ? (float)(mm + 1) : (float)mm) * dl;
    float xAEC = 0, xBD = 0, xtB = 0, xtD = 0, t2BD = 0, tB = 0, tD = 0;
    int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
    int nAx = 0, nBx = 0, nCx = 0, nDx = 0, nEx = 0;
    const float tAB = (float)((double)(-W) * 0.25 - (double)L * 0.5);
    const float tBC = (float)((double)(-W) * 0.25 + (double)L * 0.5);
    const float tCD = -tBC, tDE = -tAB;
    for (int i = 0; i < nsc - nSx; ++i) {
      const float tnow = ct[i + nSx];
      if (tnow > tAB) break;
      if (tnow >= tStart) { nA++; xAEC = (xAEC + cache[i]) - mean; }
      nAx++;
    }
    for (int i = 0; i < nsc - nAx - nSx; ++i) {
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
    for (int i = 0; i < nsc - nAx - nBx - nSx; ++i) {
      const float tnow = ct[i + nSx + nAx + nBx];
      if (tnow > tCD) break;
      if (tnow >= tStart) { nC++; xAEC = (xAEC + cache[i + nSx + nAx + nBx]) - mean; }
      nCx++;
    }
    for (int i = 0; i < nsc - nSx - nAx - nBx - nCx; ++i) {
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
    for (int i = 0; i < nsc - nAx - nBx - nCx - nDx - nSx; ++i) {
      const float tnow = ct[i + nSx + nAx + nBx + nCx + nDx];
      if (tnow > tEnd) break;
      if (tnow >= tStart) {
        nE++;
        xAEC = (xAEC + cache[i + nSx + nAx + nBx + nCx + nDx]) - mean;
      }
      nEx++;
    }
    const float filfac = (float)(nA + nB + nC + nD + nE) / (W / deltat);
    if ((double)filfac > kFilfac) {
      const double dL = (double)L, dW = (double)W;
      const float denA = (float)(4.0 * dL * dL * (double)(float)(nA - nC + nE) +
                                 (double)(W * W * (float)(nB + nD)) +
                                 8.0 * dW * (double)(tB - tD) + 16.0 * (double)t2BD);
      const float numB = (float)(8.0 * dL * dL * (double)xAEC - 4.0 * dW * dL * (double)xBD +
                                 16.0 * dL * (double)(xtD - xtB));
      const float Hmax = -numB / denA;
      const float Hmaxh = (float)((double)Hmax / 2.0);
      const float HL = Hmax / L;
      float res = 0.0f;
      const int eA = nSx + nAx, eB = eA + nBx, eC = eB + nCx, eD = eC + nDx, eE = eD + nEx;
      for (int i = nSx; i < eA; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) + Hmaxh; res = fmaf(e, e, res); }
      for (int i = eA; i < eB; ++i)
        if (ct[i] >= tStart) {
          const float e = (float)((double)(cache[i] - mean) -
                                  (double)HL * ((double)ct[i] + dW / 4.0));
          res = fmaf(e, e, res);
        }
      for (int i = eB; i < eC; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) - Hmaxh; res = fmaf(e, e, res); }
      for (int i = eC; i < eD; ++i)
        if (ct[i] >= tStart) {
          const float e = (float)((double)(cache[i] - mean) +
                                  (double)HL * ((double)ct[i] - dW / 4.0));
          res = fmaf(e, e, res);
        }
      for (int i = eD; i < eE; ++i)
        if (ct[i] >= tStart) { const float e = (cache[i] - mean) + Hmaxh; res = fmaf(e, e, res); }
      if (res > 0.0f) {
        const float hv = mode ? fabsf(Hmax) : Hmax;
        const float sn = hv / sqrtf(res / (float)(nA + nB + nC + nD + nE - 3));
        if (sn > res_max) { res_max = sn; Lmax = L; Hmaxmax = Hmax; }
      }
    }
  }
}

// grid: (ceil(nt/epb), nq); block: kGtThreads; epb epochs per block.
__global__ __launch_bounds__(kGtThreads) void gtls_kernel(
    float *__restrict__ tlssn, float *__restrict__ tlsw, float *__restrict__ tlst0,
    float *__restrict__ tlsl, float *__restrict__ tlshmax, const float *__restrict__ lc,
    const float *__restrict__ tu, int nt, int nq, int nw, int nl, int nsc, float wmax,
    float dw, float deltat, int mode, int epb) {
  extern __shared__ float sm[];
  const int q = blockIdx.y;
  const int k0 = blockIdx.x * epb;
  const int nep = min(epb, nt - k0);
  const int nsch = nsc / 2;
  const int tid = threadIdx.x;
  // stage one scoop per epoch: [flux nsc | time nsc]
  for (int e = 0; e < nep; ++e) {
    const int k = k0 + e;
    float *cache = sm + (size_t)e * 2 * nsc;
    for (int i = tid; i < nsc; i += kGtThreads) {
      float f = 0.0f, tt = (float)kFill;
      const int s = k + i - nsch;
      if (s >= 0 && s < nt) {
        f = 0.0f + lc[(size_t)s * nq + q];
        tt = tt + (float)((double)tu[(size_t)s * nq + q] - kFill);
      }
      if (tt >= 0.0f) tt = fmaf(-deltat, (float)k, tt);
      else if (tt < 0.0f) { tt = -1000.0f; f = 0.0f; }
      cache[i] = f;
      cache[i + nsc] = tt;
    }
  }
  __syncthreads();
  float *red = sm + (size_t)epb * 2 * nsc;  // [epb][4]: sn, W, L, H of the winner
  int *redw = reinterpret_cast<int *>(red + (size_t)epb * 4);
  if (tid < epb) redw[tid] = 0x7fffffff;
  // threads: (epoch, W) with W fastest; widths beyond the block loop
  const int per_ep = max(1, kGtThreads / epb);
  const int e = tid / per_ep, lane = tid % per_ep;
  float b_sn = 0.0f, b_L = 0.0f, b_H = 0.0f;
  int b_iw = 0x7fffffff;
  if (e < nep) {
    const float *cache = sm + (size_t)e * 2 * nsc;
    for (int iw = lane; iw < nw; iw += per_ep) {
      const float W = fmaf(-dw, (float)iw, wmax);
      float rm, Lm, Hm;
      gtls_thread(cache, nsc, nl, W, deltat, mode, rm, Lm, Hm);
      if (b_iw == 0x7fffffff || rm > b_sn) { b_sn = rm; b_L = Lm; b_H = Hm; b_iw = iw; }
    }
  }
  __syncthreads();
  // per-epoch maximum, lowest iw on ties: two-pass (max, then min index)
  __shared__ float smax[kGtThreads];
  smax[tid] = (e < nep && b_iw != 0x7fffffff) ? b_sn : -INFINITY;
  __syncthreads();
  if (lane == 0 && e < nep) {
    float m = -INFINITY;
    for (int l = 0; l < per_ep; ++l) m = fmaxf(m, smax[e * per_ep + l]);
    red[e * 4 + 0] = m;
  }
  __syncthreads();
  if (e < nep && b_iw != 0x7fffffff && b_sn == red[e * 4 + 0]) atomicMin(&redw[e], b_iw);
  __syncthreads();
  if (e < nep && b_iw != 0x7fffffff && b_iw == redw[e]) {
    const int k = k0 + e;
    const size_t o = (size_t)k * nq + q;
    tlssn[o] = b_sn;
    tlst0[o] = deltat * (float)k;
    tlsw[o] = fmaf(-dw, (float)b_iw, wmax);
    tlsl[o] = b_L;
    tlshmax[o] = b_H;
  }
}


// ---------------------------------------------------------------------------
// Fast mode: the same statistic with float64 arithmetic, O(log nsc) per trial.
//
// Region membership is decided exactly as the reference decides it: the
// shifted time of row j for epoch k is ct = fl32(T_j - deltat*k) (one
// rounding, the reference's fused multiply-add) and a region ends at the
// first row whose ct exceeds the region's float32 boundary (the sequential
// scans of gtls_simple.h:199-305 break there).  When the valid stored times
// of a tile are non-decreasing, that row is found by a binary search over
// the scoop on keys filled forward over masked rows (the running maximum of
// the valid times), which gives the first valid row past the boundary for
// every region in turn.  The sums of a region are differences of float64
// prefix sums over the tile (count, x', x'^2, q, q^2, x'q with
// x' = x - x_ref and q = T - r the exact time relative to the tile's
// central epoch), moved to epoch coordinates tau = q - (deltat*k - r); the
// residual is expanded in them (y = x - mean):
//   A, E: sum (y + H/2)^2, C: sum (y - H/2)^2,
//   B: sum (y - (H/L)(tau + W/4))^2, D: sum (y + (H/L)(tau - W/4))^2.
// Region A's flux sum keeps the reference's index quirk (gtls_simple.h:
// 214-221: the flux of scoop row i with the time test of row i + nSx).
// A tile whose valid times are not non-decreasing is evaluated by the
// sequential scans in float64 (same results by definition).
constexpr int kGfThreads = 256;
constexpr int kGfUnits = 4;   // (epoch, W) units per thread

struct GfLds {
  float *T;       // stored time (-1000 + fl(tu + 1000)); < 0: masked
  float *K;       // forward-filled search key (running max of valid T, -inf before)
  float *xr;      // raw flux (0 for masked rows)
  int *V;         // prefix count of valid rows
  double *P;      // prefix sums [R+1][5]: x', x'^2, q, q^2, x'q over valid rows
};

// N searches, each for the first row j in [0, nsc) of the scoop at union row
// e with fl32(K - deltat*k) > thr[i] (>= when ge[i]), nsc if none, in
// lockstep with a fixed depth: every step tests row lo + step - 1 of each (a
// row at or past nsc counts as passing), so the N independent LDS reads of a
// step are in flight together.  top = the smallest power of two >= nsc.
template <int N>
__device__ __forceinline__ void gf_firstN(const GfLds &L, int e, int nsc, int top, float mdk,
                                          float kf, const float *thr, const bool *ge,
                                          int *out) {
  int lo[N];
#pragma unroll
  for (int i = 0; i < N; ++i) lo[i] = 0;
  for (int step = top; step > 0; step >>= 1) {
    float c[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int j = min(lo[i] + step - 1, nsc - 1);
      c[i] = fmaf(mdk, kf, L.K[e + j]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const bool pass = ge[i] ? c[i] >= thr[i] : c[i] > thr[i];
      const bool fail = (lo[i] + step <= nsc) & !pass;
      lo[i] += fail ? step : 0;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = lo[i];
}

// From one L to the next a region boundary (the first row at or after which
// every row passes fl32(K - deltat*k) > thr, a row at or past nsc passing)
// moves by a few rows, in a known direction.  For a threshold that did not
// decrease the boundary can only move up from its previous row b: the rows
// failing the test from b on are a prefix of b, b+1, ... (the test is
// monotone in the row, K being non-decreasing), so four rows tested at once
// advance b by their number of failures, and another round runs only when
// all four fail.  The four reads are issued together, unconditionally (rows
// clamped into the scoop, the bounds applied to the test): a row-by-row walk
// in both directions took 96.2 ms at the Kepler setting, this 75.5 (two rows
// at once 78.1, three 75.6, six 80.0; bounds-checked reads, which the
// compiler turns into one branch and one LDS round trip per row, 86.9).
constexpr int kGfProbe = 4;
__device__ __forceinline__ int gf_up(const GfLds &L, int e, int nsc, float mdk, float kf,
                                     float thr, int b) {
  for (;;) {
    int f = 0;
#pragma unroll
    for (int j = 0; j < kGfProbe; ++j) {
      const int r = b + j;  // rows at or past nsc pass
      const float ct = fmaf(mdk, kf, L.K[e + min(r, nsc - 1)]);  // every read issued
      f += (int)((r < nsc) & !(ct > thr));
    }
    b += f;
    if (f < kGfProbe) return b;
  }
}
// and for a boundary whose threshold did not increase (it can only move
// down): the rows passing below b are a prefix of b-1, b-2, ...
__device__ __forceinline__ int gf_down(const GfLds &L, int e, int nsc, float mdk, float kf,
                                       float thr, int b) {
  for (;;) {
    int p = 0;
#pragma unroll
    for (int j = 1; j <= kGfProbe; ++j) {
      const int r = b - j;
      const float ct = fmaf(mdk, kf, L.K[e + max(r, 0)]);
      p += (int)((r >= 0) & (ct > thr));
    }
    b -= p;
    if (p < kGfProbe) return b;
  }
}

struct GfSum {  // sums over the valid rows of a union row range, tau coordinates
  double n, x, x2, t, t2, xt;
};
// kShift = false: the epoch is the tile's reference (s = 0), and the moved
// moments equal the raw ones exactly (q - n*0 = q), so the shift is skipped
template <bool kShift = true>
__device__ __forceinline__ GfSum gf_sum(const GfLds &L, int a, int b, double s) {
  const double *pa = L.P + 5 * (size_t)a, *pb = L.P + 5 * (size_t)b;
  GfSum o;
  o.n = (double)(L.V[b] - L.V[a]);
  o.x = pb[0] - pa[0];
  o.x2 = pb[1] - pa[1];
  const double q1 = pb[2] - pa[2], q2 = pb[3] - pa[3], xq = pb[4] - pa[4];
  if constexpr (kShift) {
    o.t = q1 - o.n * s;
    o.t2 = q2 - 2.0 * s * q1 + o.n * s * s;
    o.xt = xq - s * o.x;
  } else {
    o.t = q1;
    o.t2 = q2;
    o.xt = xq;
  }
  return o;
}

// Sequential float64 scans of one (epoch, W) (the definition; used for tiles
// whose valid times are not non-decreasing).  Rows e .. e+nsc-1 of the tile.
__device__ void gf_scan64(const GfLds &L, int e, int nsc, int nl, float W, float deltat,
                          float kf, int mode, double &res_max, float &Lmax, double &Hmaxmax) {
  const float dl = mode ? (float)((double)W / 4.0 / (double)(float)nl)
                        : (float)((double)W * 0.45 / (double)(float)nl);
  const float tEnd = (float)(0.5 * (double)W), tStart = (float)(-0.5 * (double)W);
  auto ctof = [&](int i) {
    const float T = L.T[e + i];
    return T >= 0.0f ? fmaf(-deltat, kf, T) : -1000.0f;
  };
  const double dtk = (double)deltat * (double)kf;
  auto tauof = [&](int i) { return (double)L.T[e + i] - dtk; };  // exact shifted time
  auto xof = [&](int i) { return (double)L.xr[e + i]; };
  double msum = 0.0;
  int mcnt = 0;
  for (int m = 0; m < nsc; ++m) {
    const float c = ctof(m);
    if (c >= tStart && c <= tEnd) { msum += xof(m); mcnt++; }
  }
  const double mean = mcnt > 0 ? msum / (double)mcnt : 0.0;
  int nSx = 0;
  for (int i = 0; i < nsc; ++i) { if (ctof(i) > tStart) break; nSx++; }
  res_max = 0.0; Hmaxmax = 0.0; Lmax = 0.0f;
  const double dW = (double)W;
  for (int mm = 0; mm < nl; ++mm) {
    const float Lf = (mode ? (float)(mm + 1) : (float)mm) * dl;
    const double dL = (double)Lf;
    double xAEC = 0, xBD = 0, xtB = 0, xtD = 0, t2BD = 0, tB = 0, tD = 0;
    int nA = 0, nB = 0, nC = 0, nD = 0, nE = 0;
    int nAx = 0, nBx = 0, nCx = 0, nDx = 0, nEx = 0;
    const float tAB = (float)((double)(-W) * 0.25 - (double)Lf * 0.5);
    const float tBC = (float)((double)(-W) * 0.25 + (double)Lf * 0.5);
    const float tCD = -tBC, tDE = -tAB;
    for (int i = 0; i < nsc - nSx; ++i) {
      const float c = ctof(i + nSx);
      if (c > tAB) break;
      if (c >= tStart) { nA++; xAEC += xof(i) - mean; }
      nAx++;
    }
    for (int i = nSx + nAx; i < nsc; ++i) {
      const float c = ctof(i);
      if (c > tBC) break;
      if (c >= tStart) {
        const double y = xof(i) - mean, tt = tauof(i);
        nB++; xBD += y; tB += tt; xtB += y * tt; t2BD += tt * tt;
      }
      nBx++;
    }
    for (int i = nSx + nAx + nBx; i < nsc; ++i) {
      const float c = ctof(i);
      if (c > tCD) break;
      if (c >= tStart) { nC++; xAEC += xof(i) - mean; }
      nCx++;
    }
    for (int i = nSx + nAx + nBx + nCx; i < nsc; ++i) {
      const float c = ctof(i);
      if (c > tDE) break;
      if (c >= tStart) {
        const double y = xof(i) - mean, tt = tauof(i);
        nD++; xBD += y; tD += tt; xtD += y * tt; t2BD += tt * tt;
      }
      nDx++;
    }
    for (int i = nSx + nAx + nBx + nCx + nDx; i < nsc; ++i) {
      const float c = ctof(i);
      if (c > tEnd) break;
      if (c >= tStart) { nE++; xAEC += xof(i) - mean; }
      nEx++;
    }
    const int n = nA + nB + nC + nD + nE;
    const float filfac = (float)n / (W / deltat);
    if (!((double)filfac > kFilfac)) continue;
    const double denA = 4.0 * dL * dL * (double)(nA - nC + nE) + dW * dW * (double)(nB + nD) +
                        8.0 * dW * (tB - tD) + 16.0 * t2BD;
    const double numB = 8.0 * dL * dL * xAEC - 4.0 * dW * dL * xBD + 16.0 * dL * (xtD - xtB);
    const double H = -numB / denA, Hh = H / 2.0, HL = H / dL;
    double res = 0.0;
    const int eA = nSx + nAx, eB = eA + nBx, eC = eB + nCx, eD = eC + nDx, eE = eD + nEx;
    for (int i = nSx; i < eE; ++i) {
      if (!(ctof(i) >= tStart)) continue;
      const double y = xof(i) - mean;
      double r;
      if (i < eA) r = y + Hh;
      else if (i < eB) r = y - HL * (tauof(i) + dW / 4.0);
      else if (i < eC) r = y - Hh;
      else if (i < eD) r = y + HL * (tauof(i) - dW / 4.0);
      else r = y + Hh;
      res += r * r;
    }
    if (res > 0.0) {
      const double hv = mode ? fabs(H) : H;
      const double sn = hv / sqrt(res / (double)(n - 3));
      if (sn > res_max) { res_max = sn; Lmax = Lf; Hmaxmax = H; }
    }
  }
}

// grid: (ceil(nt/E), nq); block kGfThreads; E epochs per block (E*nw <=
// kGfThreads*kGfUnits); dynamic LDS sized by gf_lds_bytes(E + nsc - 1).
__global__ __launch_bounds__(kGfThreads, 4) void gtls_fast_kernel(
    float *__restrict__ tlssn, float *__restrict__ tlsw, float *__restrict__ tlst0,
    float *__restrict__ tlsl, float *__restrict__ tlshmax, const float *__restrict__ lc,
    const float *__restrict__ tu, int nt, int nq, int nw, int nl, int nsc, float wmax,
    float dw, float deltat, int mode, int E) {
  extern __shared__ __align__(16) unsigned char gf_sm[];
  __shared__ unsigned long long s_best[64];
  __shared__ int s_win[64];
  __shared__ int s_bad;
  __shared__ double s_tot[kGfThreads / kWave][5];
  __shared__ int s_vtot[kGfThreads / kWave];
  __shared__ float s_kmax[kGfThreads / kWave];
  __shared__ double s_xref;
  const int q = blockIdx.y;
  const int k0 = blockIdx.x * E;
  const int nep = min(E, nt - k0);
  const int nsch = nsc / 2;
  const int R = E + nsc - 1;  // union of the tile's scoops: rows k0 - nsch + r
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  GfLds L;
  L.P = reinterpret_cast<double *>(gf_sm);
  L.V = reinterpret_cast<int *>(L.P + 5 * (size_t)(R + 1));
  L.T = reinterpret_cast<float *>(L.V + (R + 1));
  L.K = L.T + R;
  L.xr = L.K + R;
  // reference time of the prefix sums: the tile's central epoch
  const double rref = (double)deltat * (double)(k0 + (nep - 1) / 2);
  if (tid == 0) {
    s_bad = 0;
    s_xref = 0.0;
  }
  if (tid < 64) { s_best[tid] = 0ull; s_win[tid] = 0x7fffffff; }
  // ---- stage the union: each thread owns a contiguous run of rows
  const int per = (R + kGfThreads - 1) / kGfThreads;
  const int r0 = tid * per, r1 = min(R, r0 + per);
  for (int r = r0; r < r1; ++r) {
    const int srow = k0 - nsch + r;
    float f = 0.0f, T = (float)kFill;
    if (srow >= 0 && srow < nt) {
      f = lc[(size_t)srow * nq + q];
      T = T + (float)((double)tu[(size_t)srow * nq + q] - kFill);
    }
    if (!(T >= 0.0f)) { T = -1000.0f; f = 0.0f; }
    L.T[r] = T;
    L.xr[r] = f;
  }
  __syncthreads();
  // x_ref: the first valid row's flux (any constant cancels in the statistic)
  if (tid == 0) {
    for (int r = 0; r < R; ++r)
      if (L.T[r] >= 0.0f) { s_xref = (double)L.xr[r]; break; }
  }
  __syncthreads();
  const double xref = s_xref;
  // ---- per-thread totals, then a block scan (sums and the running max)
  double acc[5] = {0, 0, 0, 0, 0};
  int vcnt = 0;
  float kmx = -INFINITY;
  bool mono = true;
  for (int r = r0; r < r1; ++r) {
    const float T = L.T[r];
    if (T >= 0.0f) {
      const double xp = (double)L.xr[r] - xref, qq = (double)T - rref;
      acc[0] += xp; acc[1] += xp * xp; acc[2] += qq; acc[3] += qq * qq; acc[4] += xp * qq;
      vcnt++;
      mono &= !(T < kmx);
      kmx = fmaxf(kmx, T);
    }
  }
  // wave inclusive scans (shuffles), then the waves' totals
  double inc[5];
  int vinc = vcnt;
  float kinc = kmx;
#pragma unroll
  for (int v = 0; v < 5; ++v) inc[v] = acc[v];
  for (int o = 1; o < kWave; o <<= 1) {
    double t5[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) t5[v] = __shfl_up(inc[v], o);
    const int tv = __shfl_up(vinc, o);
    const float tk = __shfl_up(kinc, o);
    if (lane >= o) {
#pragma unroll
      for (int v = 0; v < 5; ++v) inc[v] += t5[v];
      vinc += tv;
      kinc = fmaxf(kinc, tk);
    }
  }
  if (lane == kWave - 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) s_tot[wv][v] = inc[v];
    s_vtot[wv] = vinc;
    s_kmax[wv] = kinc;
  }
  __syncthreads();
  double run[5];
  int vrun = vinc - vcnt;
  float krun = __shfl_up(kinc, 1);
  if (lane == 0) krun = -INFINITY;
#pragma unroll
  for (int v = 0; v < 5; ++v) run[v] = inc[v] - acc[v];
  for (int w = 0; w < wv; ++w) {
#pragma unroll
    for (int v = 0; v < 5; ++v) run[v] += s_tot[w][v];
    vrun += s_vtot[w];
    krun = fmaxf(krun, s_kmax[w]);
  }
  // the valid times are non-decreasing iff no valid row lies below the
  // running maximum before it (within the thread's run: checked above)
  {
    bool ok = mono;
    for (int r = r0; r < r1; ++r)
      if (L.T[r] >= 0.0f) { ok &= !(L.T[r] < krun); break; }
    if (!ok) atomicOr(&s_bad, 1);
  }
  for (int r = r0; r < r1; ++r) {
    double *pr = L.P + 5 * (size_t)r;
#pragma unroll
    for (int v = 0; v < 5; ++v) pr[v] = run[v];
    L.V[r] = vrun;
    const float T = L.T[r];
    if (T >= 0.0f) {
      const double xp = (double)L.xr[r] - xref, qq = (double)T - rref;
      run[0] += xp; run[1] += xp * xp; run[2] += qq; run[3] += qq * qq; run[4] += xp * qq;
      vrun++;
      krun = fmaxf(krun, T);
    }
    L.K[r] = krun;
  }
  if (r1 == R && r0 < r1) {
    double *pr = L.P + 5 * (size_t)R;
#pragma unroll
    for (int v = 0; v < 5; ++v) pr[v] = run[v];
    L.V[R] = vrun;
  }
  if (R <= 0 && tid == 0) {
    for (int v = 0; v < 5; ++v) L.P[v] = 0.0;
    L.V[0] = 0;
  }
  __syncthreads();
  const bool bad = s_bad != 0;
  int top = 1;
  while (top < nsc) top <<= 1;
  // ---- (epoch, W) units: W fastest
  const int nunits = nep * nw;
  double u_sn[kGfUnits], u_H[kGfUnits];
  float u_L[kGfUnits];
  int u_id[kGfUnits];
#pragma unroll
  for (int u = 0; u < kGfUnits; ++u) {
    u_sn[u] = -1.0; u_H[u] = 0.0; u_L[u] = 0.0f; u_id[u] = -1;
    const int id = tid + u * kGfThreads;
    if (id >= nunits) continue;
    const int e = id / nw, iw = id % nw;
    const int k = k0 + e;
    const float kf = (float)k, mdk = -deltat;
    const float W = fmaf(-dw, (float)iw, wmax);
    const double sft = (double)deltat * (double)k - rref;
    double best = 0.0, bH = 0.0;
    float bL = 0.0f;
    if (bad) {
      gf_scan64(L, e, nsc, nl, W, deltat, kf, mode, best, bL, bH);
    } else {
      double bh2 = 0.0, bres = 1.0;  // hv^2 (n-3) L^2 and res L^2 of the best trial
      const float dl = mode ? (float)((double)W / 4.0 / (double)(float)nl)
                            : (float)((double)W * 0.45 / (double)(float)nl);
      const float tEnd = (float)(0.5 * (double)W), tStart = (float)(-0.5 * (double)W);
      // the region thresholds of step mm (the L grid's mm-th L)
      auto thresholds = [&](int mm, float &Lf, float *thr) {
        Lf = (mode ? (float)(mm + 1) : (float)mm) * dl;
        thr[0] = (float)((double)(-W) * 0.25 - (double)Lf * 0.5);  // tAB
        thr[1] = (float)((double)(-W) * 0.25 + (double)Lf * 0.5);  // tBC
        thr[2] = -thr[1];                                          // tCD
        thr[3] = -thr[0];                                          // tDE
      };
      // the window's ends (nSx: > tStart, mlo: >= tStart, eE: > tEnd) and
      // the first L's four region ends, seven searches in lockstep
      int bnd0[4], nSx, mlo, eE;
      {
        float th[7], Lf0;
        th[0] = tStart; th[1] = tStart; th[2] = tEnd;
        thresholds(0, Lf0, th + 3);
        const bool ge[7] = {false, true, false, false, false, false, false};
        int r[7];
        gf_firstN<7>(L, e, nsc, top, mdk, kf, th, ge, r);
        nSx = r[0]; mlo = r[1]; eE = r[2];
        bnd0[0] = r[3]; bnd0[1] = r[4]; bnd0[2] = r[5]; bnd0[3] = r[6];
      }
      // mean over the closed window (valid rows with tStart <= ct <= tEnd)
      const GfSum M = gf_sum(L, e + mlo, e + eE, sft);
      const double meanp = M.n > 0.0 ? M.x / M.n : -xref;  // x' units (mean 0 if empty)
      const GfSum Wn = gf_sum(L, e + nSx, e + eE, sft);   // regions A..E together
      const double n = Wn.n;
      const float filfac = (float)(int)n / (W / deltat);
      if ((double)filfac > kFilfac) {
        const double dW = (double)W;
        // y = x' - mean' on the window: sum y^2 (the residual's constant)
        const double Y2 = Wn.x2 - 2.0 * meanp * Wn.x + n * meanp * meanp;
        // the L sweep, with the moments' shift to epoch coordinates elided
        // when the epoch is the tile's reference (one epoch per tile)
        const double px0 = L.P[5 * (size_t)e];  // the scoop's first prefix row
        const int v0 = L.V[e];
        auto sweep = [&](auto shift) {
        constexpr bool kSh = decltype(shift)::value;
        int bnd[4] = {bnd0[0], bnd0[1], bnd0[2], bnd0[3]};
        for (int mm = 0; mm < nl; ++mm) {
          float Lf, thr[4];
          thresholds(mm, Lf, thr);
          const double dL = (double)Lf;
          if (mm > 0) {
            // L grows with mm: tAB and tCD fall (their rows move down), tBC
            // and tDE rise (up)
            bnd[0] = gf_down(L, e, nsc, mdk, kf, thr[0], bnd[0]);
            bnd[1] = gf_up(L, e, nsc, mdk, kf, thr[1], bnd[1]);
            bnd[2] = gf_down(L, e, nsc, mdk, kf, thr[2], bnd[2]);
            bnd[3] = gf_up(L, e, nsc, mdk, kf, thr[3], bnd[3]);
          }
          const int eA = bnd[0], eB = bnd[1], eC = bnd[2], eD = bnd[3];
          const GfSum A = gf_sum<kSh>(L, e + nSx, e + eA, sft);
          const GfSum B = gf_sum<kSh>(L, e + eA, e + eB, sft);
          const GfSum Cc = gf_sum<kSh>(L, e + eB, e + eC, sft);
          const GfSum D = gf_sum<kSh>(L, e + eC, e + eD, sft);
          const GfSum Ee = gf_sum<kSh>(L, e + eD, e + eE, sft);
          // region A's flux sum with the index quirk: the flux of scoop rows
          // [0, eA - nSx) for the time tests of rows [nSx, eA)
          double sAq;
          const int cnt = eA - nSx;
          // the quirk rows' prefix row end, read with the region rows (before
          // the branch: one LDS round trip fewer per trial)
          const double qx = L.P[5 * (size_t)(e + cnt)];
          const int qv = L.V[e + cnt];
          if ((int)A.n == cnt) {
            // every tested row is valid: a prefix difference over the quirk
            // rows (masked rows there carry flux 0, i.e. x' = -x_ref)
            sAq = (qx - px0) - A.n * meanp + xref * (double)(qv - v0 - cnt);
          } else {
            double sr = 0.0;
            for (int i = 0; i < cnt; ++i)
              if (L.T[e + i + nSx] >= 0.0f) sr += (double)L.xr[e + i];
            sAq = sr - A.n * (meanp + xref);
          }
          const double yB = B.x - B.n * meanp, yD = D.x - D.n * meanp;
          const double yA = A.x - A.n * meanp, yC = Cc.x - Cc.n * meanp,
                       yE = Ee.x - Ee.n * meanp;
          const double ytB = B.xt - meanp * B.t, ytD = D.xt - meanp * D.t;
          const double xAEC = sAq + yC + yE;
          const double xBD = yB + yD;
          const double denA = 4.0 * dL * dL * (A.n - Cc.n + Ee.n) + dW * dW * (B.n + D.n) +
                              8.0 * dW * (B.t - D.t) + 16.0 * (B.t2 + D.t2);
          const double numB = 8.0 * dL * dL * xAEC - 4.0 * dW * dL * xBD +
                              16.0 * dL * (ytD - ytB);
          // the residual with h = H/2 and a = H/L,
          //   res = Y2 + 2h(yA - yC + yE) + h^2 (nA + nC + nE)
          //         - 2a(ytB + q4 yB) + a^2 (B.t2 + 2 q4 B.t + q4^2 B.n)
          //         + 2a(ytD - q4 yD) + a^2 (D.t2 - 2 q4 D.t + q4^2 D.n),
          // is formed times L^2 (no division by L; L = 0 empties B and D, so
          // that trial's scaled residual is 0, no trial, as its residual is
          // NaN); S/N = hv / sqrt(res/(n-3)) > 0 is compared with the best's
          // as hv^2 (n-3) L^2 / (res L^2), cross-multiplied (both positive),
          // and formed for the best trial only
          const double H = -numB / denA, q4 = dW / 4.0, L2 = dL * dL;
          const double rs =
              L2 * (Y2 + H * (yA - yC + yE) + 0.25 * H * H * (A.n + Cc.n + Ee.n)) +
              2.0 * H * dL * ((ytD - q4 * yD) - (ytB + q4 * yB)) +
              H * H * ((B.t2 + 2.0 * q4 * B.t + q4 * q4 * B.n) +
                       (D.t2 - 2.0 * q4 * D.t + q4 * q4 * D.n));
          const double hv = mode ? fabs(H) : H;
          if (rs > 0.0 && hv > 0.0) {
            const double h2 = hv * hv * (n - 3.0) * L2;
            if (h2 * bres > bh2 * rs) { bh2 = h2; bres = rs; bL = Lf; bH = H; }
          }
        }
        };
        if (sft == 0.0) sweep(std::false_type{}); else sweep(std::true_type{});
      }
      if (bh2 > 0.0) best = sqrt(bh2 / bres);
    }
    u_sn[u] = best; u_H[u] = bH; u_L[u] = bL; u_id[u] = id;
  }
  // ---- per-epoch maximum, lowest W index on ties
#pragma unroll
  for (int u = 0; u < kGfUnits; ++u)
    if (u_id[u] >= 0)
      atomicMax(&s_best[u_id[u] / nw], (unsigned long long)__double_as_longlong(u_sn[u]));
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kGfUnits; ++u)
    if (u_id[u] >= 0) {
      const int e = u_id[u] / nw;
      if ((unsigned long long)__double_as_longlong(u_sn[u]) == s_best[e])
        atomicMin(&s_win[e], u_id[u] % nw);
    }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kGfUnits; ++u)
    if (u_id[u] >= 0) {
      const int e = u_id[u] / nw, iw = u_id[u] % nw;
      if (iw == s_win[e]) {
        const int k = k0 + e;
        const size_t o = (size_t)k * nq + q;
        tlssn[o] = (float)u_sn[u];
        tlst0[o] = deltat * (float)k;
        tlsw[o] = fmaf(-dw, (float)iw, wmax);
        tlsl[o] = u_L[u];
        tlshmax[o] = (float)u_H[u];
      }
    }
}

// dynamic LDS of a fast-mode tile of R union rows
static size_t gf_lds_bytes(int R) {
  return (size_t)(R + 1) * (5 * sizeof(double) + sizeof(int)) + (size_t)R * 3 * sizeof(float);
}

}  // namespace gtrap

extern "C" int gtrap_gtls(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
                          const float *lc, const float *tu, int32_t nt, int32_t nq, int32_t nw,
                          int32_t nl, int32_t nsc, float t0min, float dt, float wmax, float dw,
                          float deltat, int mode, void *stream) {
  (void)t0min;
  (void)dt;  // accepted and unused, as in include/gtls_simple.h:1
  GTRAP_REQUIRE(tlssn && tlsw && tlst0 && tlsl && tlshmax && lc && tu, "gtrap_gtls: null pointer");
  const int parity = (mode & GTRAP_GTLS_PARITY) != 0;
  mode &= ~GTRAP_GTLS_PARITY;
  GTRAP_REQUIRE(mode == GTRAP_GTLS_NORMAL || mode == GTRAP_GTLS_ABSOLUTE,
                "gtrap_gtls: mode must be 0 (normal) or 1 (absolute), optionally | "
                "GTRAP_GTLS_PARITY");
  GTRAP_REQUIRE(nw >= 1 && nw <= 1024 && (nw & (nw - 1)) == 0,
                "gtrap_gtls: nw (block[0]) must be a power of two <= 1024 "
                "(include/gtls_simple.h:383-390)");
  GTRAP_REQUIRE(nt >= 1 && nq >= 1 && nq <= 65535 && nl >= 1 && nsc >= 1,
                "gtrap_gtls: bad sizes");
  if (!parity) {
    int E = (gtrap::kGfThreads * gtrap::kGfUnits) / nw;
    if (E > 64) E = 64;
    if (E < 1) E = 1;
    while (E > 1 && gtrap::gf_lds_bytes(E + nsc - 1) > 64 * 1024) E /= 2;
    const size_t shm = gtrap::gf_lds_bytes(E + nsc - 1);
    GTRAP_REQUIRE(shm <= 150 * 1024,
                  "gtrap_gtls: nsc = %d too large for the fast mode's LDS tile (at most "
                  "about 2700 rows); use GTRAP_GTLS_PARITY", nsc);
    dim3 grid((nt + E - 1) / E, nq);
    hipLaunchKernelGGL(gtrap::gtls_fast_kernel, grid, dim3(gtrap::kGfThreads), shm,
                       gtrap::as_stream(stream), tlssn, tlsw, tlst0, tlsl, tlshmax, lc, tu, nt,
                       nq, nw, nl, nsc, wmax, dw, deltat, mode, E);
    return gtrap::check_launch("gtls_fast_kernel");
  }
  int epb = gtrap::kGtThreads / (nw < gtrap::kGtThreads ? nw : gtrap::kGtThreads);
  const size_t per_ep = (size_t)2 * nsc * sizeof(float) + 4 * sizeof(float) + sizeof(int);
  while (epb > 1 && per_ep * epb > 64 * 1024) epb /= 2;
  GTRAP_REQUIRE(per_ep * epb <= 150 * 1024, "gtrap_gtls: nsc too large for LDS");
  const size_t shm = per_ep * epb;
  dim3 grid((nt + epb - 1) / epb, nq);
  hipLaunchKernelGGL(gtrap::gtls_kernel, grid, dim3(gtrap::kGtThreads), shm,
                     gtrap::as_stream(stream), tlssn, tlsw, tlst0, tlsl, tlshmax, lc, tu, nt, nq,
                     nw, nl, nsc, wmax, dw, deltat, mode, epb);
  return gtrap::check_launch("gtls_kernel");
}
