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

}  // namespace gtrap

extern "C" int gtrap_gtls(float *tlssn, float *tlsw, float *tlst0, float *tlsl, float *tlshmax,
                          const float *lc, const float *tu, int32_t nt, int32_t nq, int32_t nw,
                          int32_t nl, int32_t nsc, float t0min, float dt, float wmax, float dw,
                          float deltat, int mode, void *stream) {
  (void)t0min;
  (void)dt;  // accepted and unused, as in include/gtls_simple.h:1
  GTRAP_REQUIRE(tlssn && tlsw && tlst0 && tlsl && tlshmax && lc && tu, "gtrap_gtls: null pointer");
  GTRAP_REQUIRE(mode == GTRAP_GTLS_NORMAL || mode == GTRAP_GTLS_ABSOLUTE,
                "gtrap_gtls: mode must be 0 (normal) or 1 (absolute)");
  GTRAP_REQUIRE(nw >= 1 && nw <= 1024 && (nw & (nw - 1)) == 0,
                "gtrap_gtls: nw (block[0]) must be a power of two <= 1024 "
                "(include/gtls_simple.h:383-390)");
  GTRAP_REQUIRE(nt >= 1 && nq >= 1 && nq <= 65535 && nl >= 1 && nsc >= 1,
                "gtrap_gtls: bad sizes");
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
