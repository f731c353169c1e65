// BLS spectrum statistics (getstat) for gfx950.
//
// Reference semantics: gtrap/getstat.py:14-123.  The reference reduces each
// 1024-wide patch of the spectrum on the GPU (max, sum, sum of squares, index
// of the max; float32 tree, getstat.py:26-54) and combines the patches on the
// host (getstat.py:100-123):  the first patch holding the maximum gives idx,
// Pest = 1/(idx*df + fmin), ave = sum/nf, var = sumsq/nf - ave^2,
// sde = (max - ave)/sqrt(var), phases at idx.  Only npatch = floor(nf/patch)
// patches are reduced (getstat.py:73) while ave/var still divide by nf.
//
// Here one workgroup reduces one curve in float64 with a fixed tree and a
// lowest-index tie-break for the argmax, and writes the combined statistics:
// no host round trip and no race on ties (getstat.py:46-54 writes from every
// thread whose value equals the max).
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kStatThreads = 256;

__global__ __launch_bounds__(kStatThreads) void blsstats_kernel(
    const float *__restrict__ blsp, const uint32_t *__restrict__ phase1,
    const uint32_t *__restrict__ phase2, uint32_t nf, uint32_t nused, double df, double fmin,
    double *pest, double *sde, uint32_t *pb1, uint32_t *pb2, float *maxval) {
  const uint32_t q = blockIdx.x;
  const float *b = blsp + (size_t)nf * q;
  float mx = -INFINITY;
  uint32_t mi = 0xFFFFFFFFu;
  double s1 = 0.0, s2 = 0.0;
  for (uint32_t i = threadIdx.x; i < nused; i += kStatThreads) {
    const float v = b[i];
    if (v > mx || (v == mx && i < mi)) { mx = v; mi = i; }
    s1 += (double)v;
    s2 += (double)v * (double)v;
  }
  __shared__ float smx[kStatThreads];
  __shared__ uint32_t smi[kStatThreads];
  __shared__ double ss1[kStatThreads], ss2[kStatThreads];
  const int t = threadIdx.x;
  smx[t] = mx; smi[t] = mi; ss1[t] = s1; ss2[t] = s2;
  __syncthreads();
  for (int s = kStatThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      const float v = smx[t + s];
      const uint32_t j = smi[t + s];
      if (v > smx[t] || (v == smx[t] && j < smi[t])) { smx[t] = v; smi[t] = j; }
      ss1[t] += ss1[t + s];
      ss2[t] += ss2[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    const uint32_t idx = smi[0];
    const double ave = ss1[0] / (double)nf;
    const double var = ss2[0] / (double)nf - ave * ave;
    if (pest) pest[q] = 1.0 / ((double)idx * df + fmin);
    if (sde) sde[q] = ((double)smx[0] - ave) / sqrt(var);
    if (maxval) maxval[q] = smx[0];
    const size_t o = (size_t)nf * q + idx;
    if (pb1) pb1[q] = (phase1 && idx != 0xFFFFFFFFu) ? phase1[o] : 0xFFFFFFFFu;
    if (pb2) pb2[q] = (phase2 && idx != 0xFFFFFFFFu) ? phase2[o] : 0xFFFFFFFFu;
  }
}

}  // namespace gtrap

extern "C" int gtrap_blsstats(const float *blsp, const uint32_t *phase1, const uint32_t *phase2,
                              uint32_t nf, uint32_t nq, uint32_t patch, double df, double fmin,
                              double *pest, double *sde, uint32_t *phasebest1,
                              uint32_t *phasebest2, float *maxval, void *stream) {
  GTRAP_REQUIRE(blsp, "gtrap_blsstats: null blsp");
  GTRAP_REQUIRE(patch >= 1, "gtrap_blsstats: patch must be >= 1");
  GTRAP_REQUIRE(nq >= 1, "gtrap_blsstats: nq must be >= 1");
  const uint32_t npatch = nf / patch;
  GTRAP_REQUIRE(npatch >= 1,
                "gtrap_blsstats: nf < patch leaves no patch (gtrap/getstat.py:73)");
  hipLaunchKernelGGL(gtrap::blsstats_kernel, dim3(nq), dim3(gtrap::kStatThreads), 0,
                     gtrap::as_stream(stream), blsp, phase1, phase2, nf, npatch * patch, df,
                     fmin, pest, sde, phasebest1, phasebest2, maxval);
  return gtrap::check_launch("blsstats_kernel");
}
