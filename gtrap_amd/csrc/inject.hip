// Synthetic injection for mock light curves.
//
// gtrap_gamma: Gamma(a, rate b) deviates, one per thread (gtrap/ggenmock.py:
// 18-79: Marsaglia-Tsang with normal proposals, the u^(1/a) boost for
// a <= 1, at most MAXTRY = 10000 proposals).  Random words come from
// Philox4x32-10 (Salmon et al., SC 2011), a counter-based generator: word
// block (id, proposal) under key = seed, so a deviate depends only on (seed,
// id) and any launch shape gives the same array.  Normals by Box-Muller,
// uniforms on (0, 1].  Arithmetic in float64, stored float32.  A slot whose
// proposals all fail is NaN and counted in *nfail.
//
// gtrap_inject_trapezoid: subtract (mode 0) or multiply in (mode 1) the dip
// of the trapezoid model of gtrap/tls.py:2-15 shifted to zero outside the
// transit (depth on the flat bottom |d| < W/4 - L/2, linear ingress/egress
// to |d| = W/4 + L/2), d the offset from the nearest transit centre
// (t0 + k P; P <= 0: one event at t0), on the interleaved (n, nq) image;
// masked samples (tu < 0) are left alone.  One thread per image element,
// consecutive lanes on consecutive curves.
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr uint32_t kPhM0 = 0xD2511F53u, kPhM1 = 0xCD9E8D57u;
constexpr uint32_t kPhW0 = 0x9E3779B9u, kPhW1 = 0xBB67AE85u;

struct Ph4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ Ph4 philox4x32_10(Ph4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)kPhM0 * c.x, p1 = (uint64_t)kPhM1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = Ph4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += kPhW0;
    k1 += kPhW1;
  }
  return c;
}

__device__ __forceinline__ double ph_unit(uint32_t w) { return ((double)w + 1.0) * 0x1p-32; }

constexpr int kGammaMaxTry = 10000;

__global__ void inject_zero_i32(int32_t *p) { *p = 0; }

__global__ void gamma_kernel(float *__restrict__ x, int64_t count, double a_in, double b_in,
                             uint32_t k0, uint32_t k1, int32_t *__restrict__ nfail) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double a = a_in > 1.0 ? a_in : a_in + 1.0;
  const double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < count; id += stride) {
    float out = __int_as_float(0x7fc00000);
    for (int k = 0; k < kGammaMaxTry; ++k) {
      const Ph4 w = philox4x32_10(Ph4{(uint32_t)id, (uint32_t)(id >> 32), (uint32_t)k, 0u}, k0, k1);
      const double y = sqrt(-2.0 * log(ph_unit(w.x))) * cos(6.283185307179586 * ph_unit(w.y));
      double v = 1.0 + c * y;
      if (v <= 0.0) continue;
      v = v * v * v;
      const double u = ph_unit(w.z);
      if (u < 1.0 - 0.0331 * (y * y) * (y * y) ||
          log(u) < 0.5 * y * y + d * (1.0 - v + log(v))) {
        double g = d * v;
        if (a > a_in) g *= pow(ph_unit(w.w), 1.0 / a_in);
        out = (float)(g / b_in);
        break;
      }
    }
    if (!(out == out)) atomicAdd(nfail, 1);
    x[id] = out;
  }
}

__global__ void inject_kernel(float *__restrict__ lc, const float *__restrict__ tu, int32_t n,
                              int32_t nq, const double *__restrict__ t0,
                              const double *__restrict__ period, const double *__restrict__ depth,
                              const double *__restrict__ width, const double *__restrict__ ell,
                              int mode) {
  const int64_t total = (int64_t)n * nq;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const float tf = tu[e];
    if (tf < 0.0f) continue;
    const int q = (int)(e % nq);
    const double P = period[q], L = ell[q];
    double dd = (double)tf - t0[q];
    if (P > 0.0) {
      double m = fmod(dd + 0.5 * P, P);
      if (m < 0.0) m += P;
      dd = m - 0.5 * P;
    }
    const double a = fabs(dd), q4 = width[q] / 4.0, h = L / 2.0;
    double p = 0.0;
    if (a < q4 - h) p = depth[q];
    else if (a >= q4 - h && a < q4 + h) p = depth[q] * (q4 + h - a) / L;
    if (p != 0.0)
      lc[e] = mode ? (float)((double)lc[e] * (1.0 - p)) : (float)((double)lc[e] - p);
  }
}

}  // namespace gtrap

extern "C" int gtrap_gamma(float *x, int64_t count, double a, double b, uint64_t seed,
                           int32_t *nfail, void *stream) {
  using namespace gtrap;
  GTRAP_REQUIRE(x && nfail, "gtrap_gamma: null pointer");
  GTRAP_REQUIRE(count >= 0 && a > 0.0 && b > 0.0, "gtrap_gamma: need count >= 0, a > 0, b > 0");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(inject_zero_i32, dim3(1), dim3(1), 0, st, nfail);
  if (count > 0) {
    const int64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(gamma_kernel, dim3((unsigned)(blocks < 16384 ? blocks : 16384)), dim3(256),
                       0, st, x, count, a, b, (uint32_t)seed, (uint32_t)(seed >> 32), nfail);
  }
  return check_launch("gamma_kernel");
}

extern "C" int gtrap_inject_trapezoid(float *lc, const float *tu, int32_t n, int32_t nq,
                                      const double *t0, const double *period,
                                      const double *depth, const double *width,
                                      const double *ell, int mode, void *stream) {
  using namespace gtrap;
  GTRAP_REQUIRE(lc && tu && t0 && period && depth && width && ell,
                "gtrap_inject_trapezoid: null pointer");
  GTRAP_REQUIRE(n >= 0 && nq >= 1 && (mode == 0 || mode == 1),
                "gtrap_inject_trapezoid: need n >= 0, nq >= 1, mode 0 or 1");
  const int64_t total = (int64_t)n * nq;
  if (total == 0) return GTRAP_OK;
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(inject_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256),
                     0, as_stream(stream), lc, tu, n, nq, t0, period, depth, width, ell, mode);
  return check_launch("inject_kernel");
}
