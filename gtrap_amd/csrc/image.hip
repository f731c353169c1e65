// Image assembly: ragged light curves -> the interleaved (n, nq) image.
//
// Reference semantics: gtrap/read_keplerlc.py:23-57 (load_keplc) and
// :123-146 (throw_kepintarray), for curves already in memory.  Curve q's
// samples are t[off[q] + i], flux[...], cno[...], i < len[q]:
//   * every slot starts as flux = fillv, time = inval;
//   * sample i lands in slot j = cno[i] - cno[0] when its flux is not NaN:
//     flux[j] = (float)flux_i, time[j] = (float)(t_i - t0), t0 = t[0] or 0;
//   * gaps are slots with time < 0, ntrue = n - gaps;
//   * with H = max and L = min of the curve's n flux slots (fills included),
//     gap slots at even j get L, at odd j get H (zig-zag fill).
//
// MI355X design: the fill, statistics and zig-zag passes walk the image by
// rows (consecutive lanes on consecutive curves, 256 contiguous bytes per
// wave), so every pass is a coalesced streaming pass over HBM; per-curve
// extrema and counts are reduced in registers over a run of rows and merged
// with one atomic per thread (order-preserving integer keys for the floats).
// The scatter runs one thread per sample.  Samples whose slot falls outside
// [0, n) are skipped and counted in *nbad.  For duplicate cadence numbers
// the winning sample is unspecified (the reference keeps the later one).
#include "common.h"
#include <math.h>

namespace gtrap {

constexpr int kImThreads = 256;
constexpr int kImRows = 256;  // rows per thread in the row passes

__device__ __forceinline__ uint32_t im_key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float im_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void image_fill_kernel(float *__restrict__ lc, float *__restrict__ tu, int64_t total,
                                  float fillv, float inval, uint32_t *__restrict__ hkey,
                                  uint32_t *__restrict__ lkey, uint32_t *__restrict__ ntrue,
                                  int32_t nq, int32_t *__restrict__ nbad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    lc[e] = fillv;
    tu[e] = inval;
  }
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += stride) {
    hkey[q] = 0u;
    lkey[q] = 0xffffffffu;
    ntrue[q] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *nbad = 0;
}

__global__ void image_scatter_kernel(float *__restrict__ lc, float *__restrict__ tu,
                                     double *__restrict__ t0out, const double *__restrict__ t,
                                     const double *__restrict__ flux,
                                     const int64_t *__restrict__ cno,
                                     const int64_t *__restrict__ off,
                                     const int32_t *__restrict__ len, int32_t nq, int32_t n,
                                     int offt, int32_t *__restrict__ nbad) {
  const int q = blockIdx.y;
  const int64_t base = off[q];
  const int m = len[q];
  const double t0 = (offt && m > 0) ? t[base] : 0.0;
  const int i = blockIdx.x * kImThreads + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) t0out[q] = t0;
  if (i >= m) return;
  const int64_t j = cno[base + i] - cno[base];
  if (j < 0 || j >= n) {
    atomicAdd(nbad, 1);
    return;
  }
  const double f = flux[base + i];
  if (!(f == f)) return;
  lc[j * nq + q] = (float)f;
  tu[j * nq + q] = (float)(t[base + i] - t0);
}

__global__ void image_stats_kernel(const float *__restrict__ lc, const float *__restrict__ tu,
                                   int32_t nq, int32_t n, uint32_t *__restrict__ hkey,
                                   uint32_t *__restrict__ lkey, uint32_t *__restrict__ ntrue) {
  const int q = blockIdx.x * kImThreads + threadIdx.x;
  if (q >= nq) return;
  const int r0 = blockIdx.y * kImRows, r1 = min(n, r0 + kImRows);
  uint32_t hk = 0u, lk = 0xffffffffu, cnt = 0u;
  for (int r = r0; r < r1; ++r) {
    const size_t e = (size_t)r * nq + q;
    const uint32_t k = im_key(lc[e]);
    hk = max(hk, k);
    lk = min(lk, k);
    cnt += tu[e] < 0.0f ? 0u : 1u;
  }
  if (r1 > r0) {
    atomicMax(hkey + q, hk);
    atomicMin(lkey + q, lk);
    atomicAdd(ntrue + q, cnt);
  }
}

__global__ void image_zigzag_kernel(float *__restrict__ lc, const float *__restrict__ tu,
                                    int32_t nq, int32_t n, const uint32_t *__restrict__ hkey,
                                    const uint32_t *__restrict__ lkey) {
  const int q = blockIdx.x * kImThreads + threadIdx.x;
  if (q >= nq) return;
  const float hi = im_unkey(hkey[q]), lo = im_unkey(lkey[q]);
  const int r0 = blockIdx.y * kImRows, r1 = min(n, r0 + kImRows);
  for (int r = r0; r < r1; ++r) {
    const size_t e = (size_t)r * nq + q;
    if (tu[e] < 0.0f) lc[e] = (r & 1) ? hi : lo;
  }
}

__global__ void image_final_kernel(uint32_t *__restrict__ hkey, uint32_t *__restrict__ lkey,
                                   int32_t nq) {
  const int q = blockIdx.x * kImThreads + threadIdx.x;
  if (q >= nq) return;
  const float hi = im_unkey(hkey[q]), lo = im_unkey(lkey[q]);
  reinterpret_cast<float *>(hkey)[q] = hi;
  reinterpret_cast<float *>(lkey)[q] = lo;
}

}  // namespace gtrap

extern "C" int gtrap_image_build(float *lc, float *tu, uint32_t *ntrue, float *hfill,
                                 float *lfill, double *t0, const double *t, const double *flux,
                                 const int64_t *cno, const int64_t *off, const int32_t *len,
                                 int32_t nq, int32_t n, int32_t maxlen, float fillv, float inval,
                                 int offt, int32_t *nbad, void *stream) {
  using namespace gtrap;
  GTRAP_REQUIRE(lc && tu && ntrue && hfill && lfill && t0 && t && flux && cno && off && len &&
                    nbad,
                "gtrap_image_build: null pointer");
  GTRAP_REQUIRE(nq >= 1 && nq <= 65535 && n >= 1 && maxlen >= 0,
                "gtrap_image_build: need 1 <= nq <= 65535, n >= 1, maxlen >= 0");
  GTRAP_REQUIRE((int64_t)n <= 65535LL * kImRows && (int64_t)maxlen <= 65535LL * 65536LL,
                "gtrap_image_build: n or maxlen too large");
  hipStream_t st = as_stream(stream);
  const int64_t total = (int64_t)n * nq;
  const int64_t fb = (total + kImThreads - 1) / kImThreads;
  const unsigned gfill = (unsigned)(fb < 65536 ? (fb > 0 ? fb : 1) : 65536);
  uint32_t *hkey = reinterpret_cast<uint32_t *>(hfill), *lkey = reinterpret_cast<uint32_t *>(lfill);
  hipLaunchKernelGGL(image_fill_kernel, dim3(gfill), dim3(kImThreads), 0, st, lc, tu, total,
                     fillv, inval, hkey, lkey, ntrue, nq, nbad);
  if (maxlen > 0)
    hipLaunchKernelGGL(image_scatter_kernel,
                       dim3((unsigned)((maxlen + kImThreads - 1) / kImThreads), (unsigned)nq),
                       dim3(kImThreads), 0, st, lc, tu, t0, t, flux, cno, off, len, nq, n, offt,
                       nbad);
  else
    hipLaunchKernelGGL(image_scatter_kernel, dim3(1, (unsigned)nq), dim3(kImThreads), 0, st, lc,
                       tu, t0, t, flux, cno, off, len, nq, n, offt, nbad);
  const dim3 rows((unsigned)((nq + kImThreads - 1) / kImThreads),
                  (unsigned)((n + kImRows - 1) / kImRows));
  hipLaunchKernelGGL(image_stats_kernel, rows, dim3(kImThreads), 0, st, lc, tu, nq, n, hkey, lkey,
                     ntrue);
  hipLaunchKernelGGL(image_zigzag_kernel, rows, dim3(kImThreads), 0, st, lc, tu, nq, n, hkey,
                     lkey);
  hipLaunchKernelGGL(image_final_kernel, dim3((unsigned)((nq + kImThreads - 1) / kImThreads)),
                     dim3(kImThreads), 0, st, hkey, lkey, nq);
  return check_launch("image kernels");
}
