// Box least squares (geebls / geebls_simple) for gfx950.
//
// Reference semantics: include/geebls.h:1-167 and gtrap/geebls_simple.py:17-175.
// Per (trial frequency jf, curve q): f0 = fmin + df*jf; every sample
// i < floor(n/nthread)*nthread with tu >= 0 is folded, ph = frac(tu*f0), and
// binned into j = uint(nb*ph); bins hold the sum of (lc - offset[q]) and the
// count; the first kma bins are appended (wrap); for every start bin and
// width k = 1..kma the running sum s and count kk give
// pow = s^2 / (kk (ntrue - kk)) when k > max(kmi,1) and kk > max(kkmi,5);
// blsp = sqrt(max pow) with the start/end bins of the best box.
//
// MI355X design.  One workgroup owns one curve and a run of `fpb` trial
// frequencies, so each sample is loaded from L2 once per fpb frequencies and
// folded fpb times from registers.  Bin sums are exact 64-bit fixed-point
// LDS atomics (ds_add_u64) with a per-curve power-of-two scale, rounded once
// to float32: the result does not depend on atomic order (the reference's
// float atomics do).  The scan keeps the reference's sequential float32
// running sum per start bin (one LDS read per trial), and the argmax is a
// deterministic wave/LDS reduction with a lowest-(start, end) tie-break.
//
// With a caller workspace (gtrap_geebls_ws) one pass first copies the
// interleaved image to curve-major order and takes every curve's scale
// (max |lc - offset|) once: the binning loads are then coalesced and no
// workgroup gathers its whole curve for the scale.  Same results.
#include "common.h"
#include <math.h>

namespace gtrap {

// 512 threads and at most 52 KB of bins per workgroup: three workgroups (24
// waves) per CU keep the LDS atomics' latency covered.  On 1024 curves x
// 16384 samples x 2048 frequencies (nb 1024, kma 22): 256 threads and 64 KB
// (two workgroups of 4 waves) 72.4 ms, 512 threads and 64 KB 45.0 ms, 1024
// threads and 64 KB 40.8 ms, 512 threads and 40 KB 38.7 ms, 512 and 52 KB
// 35.3 ms.  Bin sums are order-independent, so the outputs do not depend on
// either choice.
constexpr int kBlsThreads = 512;
constexpr size_t kBlsLdsBudget = 52 * 1024;
constexpr int kBlsSpt = 16;  // samples per thread per register tile

__device__ __forceinline__ int bls_scale_exp(float maxabs, uint32_t n) {
  int e = -100;
  if (maxabs > 0.0f) frexpf(maxabs, &e);
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) <= (uint64_t)n) ++b;
  int S = 62 - e - b;
  return S > 200 ? 200 : S;
}

struct Best {
  float p;
  uint32_t i, j;
};

__device__ __forceinline__ bool better(float p, uint32_t i, uint32_t j, const Best &b) {
  return (p > b.p) || (p == b.p && (i < b.i || (i == b.i && j < b.j)));
}

constexpr int kBtTile = 64;  // transpose tile: 64 samples x 64 curves

__global__ void bls_zero_kernel(uint32_t *p, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// Interleaved (n, nq) lc (and tu unless null) to curve-major copies
// (element (i, q) to q*n + i) through a padded LDS tile, and each curve's
// max |lc - offset| over its n samples as float bits in maxbits[q] (fmaxf
// from 0: NaN samples are skipped, as in the per-workgroup scale).
__global__ __launch_bounds__(256) void bls_transpose_kernel(
    const float *__restrict__ lc, const float *__restrict__ tu, const float *__restrict__ offset,
    float *__restrict__ lct, float *__restrict__ tut, uint32_t *__restrict__ maxbits, uint32_t n,
    uint32_t nq) {
  __shared__ float sl[kBtTile][kBtTile + 1], st[kBtTile][kBtTile + 1];
  const uint32_t q0 = blockIdx.x * kBtTile, i0 = blockIdx.y * kBtTile;
  const int tx = threadIdx.x & (kBtTile - 1), ty = threadIdx.x / kBtTile;
  const uint32_t q = q0 + tx;
  const float off = (offset && q < nq) ? offset[q] : 0.0f;
  float mx = 0.0f;
  for (int r = ty; r < kBtTile; r += 256 / kBtTile) {
    const uint32_t i = i0 + r;
    float v = 0.0f, t = 0.0f;
    if (i < n && q < nq) {
      v = lc[(size_t)i * nq + q];
      mx = fmaxf(mx, fabsf(v - off));
      if (tu) t = tu[(size_t)i * nq + q];
    }
    sl[r][tx] = v;
    st[r][tx] = t;
  }
  if (q < nq) atomicMax(&maxbits[q], __float_as_uint(mx));
  __syncthreads();
  for (int r = ty; r < kBtTile; r += 256 / kBtTile) {
    const uint32_t qq = q0 + r, i = i0 + tx;
    if (qq < nq && i < n) {
      lct[(size_t)qq * n + i] = sl[tx][r];
      if (tu) tut[(size_t)qq * n + i] = st[tx][r];
    }
  }
}

// T: curve-major lc/tu (tu 1-D when SIMPLE) and the scales in maxbits
template <bool SIMPLE, bool T>
__global__ __launch_bounds__(kBlsThreads) void bls_kernel(
    float *__restrict__ blsp, uint32_t *__restrict__ phase1, uint32_t *__restrict__ phase2,
    const float *__restrict__ lc, const float *__restrict__ tu, const float *__restrict__ offset,
    const uint32_t *__restrict__ ntrue, const uint32_t *__restrict__ maxbits, uint32_t n,
    uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma, uint32_t kmi, uint32_t kkmi, float fmin,
    float df, uint32_t nused, int fpb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t nbk = nb + kma;
  // LDS: Y[fpb][nbk] int64 | C[fpb][nbk] u32 | yf[nbk] float (scan view)
  int64_t *Y = reinterpret_cast<int64_t *>(smem);
  uint32_t *Cn = reinterpret_cast<uint32_t *>(Y + (size_t)fpb * nbk);
  float *yf = reinterpret_cast<float *>(Cn + (size_t)fpb * nbk);
  __shared__ float red_p[kBlsThreads / kWave];
  __shared__ uint32_t red_i[kBlsThreads / kWave], red_j[kBlsThreads / kWave];
  __shared__ float red_max[kBlsThreads / kWave];

  const uint32_t q = blockIdx.y;
  const uint32_t jf0 = blockIdx.x * (uint32_t)fpb;
  const int nfr = (int)min((uint32_t)fpb, nf - jf0);
  const int tid = threadIdx.x;
  const float off = SIMPLE ? 0.0f : offset[q];
  const float rntrue = SIMPLE ? (float)n : (float)ntrue[q];
  const uint32_t kmi_e = kmi < 1 ? 1 : kmi;
  const uint32_t kkmi_e = kkmi < 5 ? 5 : kkmi;
  const float rnb = (float)nb;

  // per-curve fixed-point scale from max |lc - offset| over all n samples
  // sample i of curve q: curve-major or interleaved
  const auto at = [&](uint32_t i) -> size_t { return T ? (size_t)q * n + i : (size_t)i * nq + q; };
  float mx = 0.0f;
  if (!T) {
    for (uint32_t i = tid; i < n; i += kBlsThreads) mx = fmaxf(mx, fabsf(lc[at(i)] - off));
    for (int o = kWave / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((tid & (kWave - 1)) == 0) red_max[tid / kWave] = mx;
  }
  for (uint32_t e = tid; e < (uint32_t)fpb * nbk; e += kBlsThreads) {
    Y[e] = 0;
    Cn[e] = 0;
  }
  __syncthreads();
  if (T) {
    mx = __uint_as_float(maxbits[q]);
  } else {
    mx = red_max[0];
    for (int w = 1; w < kBlsThreads / kWave; ++w) mx = fmaxf(mx, red_max[w]);
  }
  const int S = bls_scale_exp(mx, n);

  float f0s[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) f0s[f] = fmaf(df, (float)(jf0 + f), fmin);

  // binning, one register tile of kBlsThreads*kBlsSpt samples at a time
  for (uint32_t base = 0; base < nused; base += kBlsThreads * kBlsSpt) {
    float tt[kBlsSpt];
    int64_t qv[kBlsSpt];
#pragma unroll
    for (int s = 0; s < kBlsSpt; ++s) {
      const uint32_t i = base + s * kBlsThreads + tid;
      float tx = -1.0f;
      int64_t v = 0;
      if (i < nused) {
        tx = SIMPLE ? tu[i] : tu[at(i)];
        v = llrint(ldexp((double)(lc[at(i)] - off), S));
      }
      tt[s] = tx;
      qv[s] = v;
    }
    for (int f = 0; f < nfr; ++f) {
      const float f0 = f0s[f];
      int64_t *Yf = Y + (size_t)f * nbk;
      uint32_t *Cf = Cn + (size_t)f * nbk;
#pragma unroll
      for (int s = 0; s < kBlsSpt; ++s) {
        const float tx = tt[s];
        const bool ok = (base + s * kBlsThreads + tid < nused) && (SIMPLE || tx >= 0.0f);
        if (ok) {
          float ph = tx * f0;
          ph = ph - (float)(int)ph;
          const float fj = rnb * ph;
          const uint32_t j = fj > 0.0f ? (uint32_t)fj : 0u;  // saturating, as cvt.rzi.u32
          if (j < nb) {
            atomicAdd(reinterpret_cast<unsigned long long *>(Yf + j),
                      (unsigned long long)qv[s]);
            atomicAdd(Cf + j, 1u);
          }
        }
      }
    }
  }
  __syncthreads();

  // wrap the first kma bins (include/geebls.h:84-95)
  for (int f = 0; f < nfr; ++f) {
    int64_t *Yf = Y + (size_t)f * nbk;
    uint32_t *Cf = Cn + (size_t)f * nbk;
    for (uint32_t k = tid; k < kma; k += kBlsThreads) {
      Yf[nb + k] = Yf[k];
      Cf[nb + k] = Cf[k];
    }
  }
  __syncthreads();

  // scan: start bins over threads, widths sequential
  for (int f = 0; f < nfr; ++f) {
    const int64_t *Yf = Y + (size_t)f * nbk;
    const uint32_t *Cf = Cn + (size_t)f * nbk;
    for (uint32_t e = tid; e < nbk; e += kBlsThreads) yf[e] = (float)ldexp((double)Yf[e], -S);
    __syncthreads();
    Best b{0.0f, 0xFFFFFFFFu, 0xFFFFFFFFu};
    for (uint32_t i = tid; i < nb; i += kBlsThreads) {
      float s = 0.0f;
      uint32_t kk = 0;
      for (uint32_t k = 1; k <= kma; ++k) {
        const uint32_t j = i + k - 1;
        s = s + yf[j];
        kk = kk + Cf[j];
        if (k > kmi_e && kk > kkmi_e) {
          const float rn1 = (float)kk;
          const float pw = (s * s) / (rn1 * (rntrue - rn1));
          if (pw > b.p) { b.p = pw; b.i = i; b.j = j; }
        }
      }
    }
    // wave reduction
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const float p2 = __shfl_xor(b.p, o);
      const uint32_t i2 = __shfl_xor(b.i, o), j2 = __shfl_xor(b.j, o);
      if (better(p2, i2, j2, b)) { b.p = p2; b.i = i2; b.j = j2; }
    }
    if ((tid & (kWave - 1)) == 0) {
      red_p[tid / kWave] = b.p;
      red_i[tid / kWave] = b.i;
      red_j[tid / kWave] = b.j;
    }
    __syncthreads();
    if (tid == 0) {
      Best r{red_p[0], red_i[0], red_j[0]};
      for (int w = 1; w < kBlsThreads / kWave; ++w)
        if (better(red_p[w], red_i[w], red_j[w], r)) r = Best{red_p[w], red_i[w], red_j[w]};
      const size_t o = (size_t)(jf0 + f) + (size_t)nf * q;
      blsp[o] = sqrtf(r.p);
      if (!SIMPLE) {
        if (phase1) phase1[o] = r.i;
        if (phase2) phase2[o] = r.j;
      }
    }
    __syncthreads();
  }
}

// Workspace of the curve-major path: per-curve scale bits, then lc and (unless
// simple) tu in curve-major order, each 256-byte aligned.
static size_t bls_ws_part(size_t bytes) { return (bytes + 255) / 256 * 256; }
static size_t bls_ws_need(uint32_t n, uint32_t nq, bool simple) {
  const size_t img = bls_ws_part((size_t)n * nq * sizeof(float));
  return bls_ws_part((size_t)nq * sizeof(uint32_t)) + img * (simple ? 1 : 2);
}

static int bls_launch(bool simple, float *blsp, uint32_t *phase1, uint32_t *phase2,
                      const float *lc, const float *tu, const float *offset,
                      const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb,
                      uint32_t kma, uint32_t kmi, uint32_t kkmi, float fmin, float df,
                      uint32_t nthread, void *ws, int64_t ws_bytes, hipStream_t st) {
  const char *fn = simple ? (ws ? "gtrap_geebls_simple_ws" : "gtrap_geebls_simple")
                          : (ws ? "gtrap_geebls_ws" : "gtrap_geebls");
  GTRAP_REQUIRE(blsp && lc && tu, "%s: null pointer", fn);
  GTRAP_REQUIRE(simple || (offset && ntrue), "%s: offset/ntrue required", fn);
  GTRAP_REQUIRE(nthread >= 1 && nthread <= 1024, "%s: nthread (block[0]) must be 1..1024", fn);
  GTRAP_REQUIRE(nb >= 1 && (nb & (nb - 1)) == 0,
                "%s: nb must be a power of two (max tree, include/geebls.h:145)", fn);
  GTRAP_REQUIRE(nb % nthread == 0,
                "%s: nb must be a multiple of nthread (bins >= floor(nb/nthread)*nthread are "
                "uninitialised in include/geebls.h:39-50)", fn);
  GTRAP_REQUIRE(kma >= 1 && kma <= nthread,
                "%s: need 1 <= kma <= nthread (wrap, include/geebls.h:84-85)", fn);
  GTRAP_REQUIRE(kma <= nb, "%s: kma must not exceed nb", fn);
  GTRAP_REQUIRE(n >= 1 && nq >= 1 && nf >= 1 && nq <= 65535, "%s: bad sizes", fn);
  if (ws) {
    if (ws_bytes < 0 || (size_t)ws_bytes < bls_ws_need(n, nq, simple)) {
      set_error("%s: workspace of %lld bytes, %zu needed", fn, (long long)ws_bytes,
                bls_ws_need(n, nq, simple));
      return GTRAP_ENOMEM;
    }
    GTRAP_REQUIRE(((uintptr_t)ws & 255) == 0, "%s: workspace must be 256-byte aligned", fn);
  }
  const uint32_t nbk = nb + kma;
  const size_t per_f = (size_t)nbk * 12;
  int fpb = 4 * (size_t)nbk < kBlsLdsBudget ? (int)((kBlsLdsBudget - 4 * (size_t)nbk) / per_f) : 1;
  if (fpb > 8) fpb = 8;
  if (fpb < 1) fpb = 1;
  GTRAP_REQUIRE(per_f * fpb + 4 * (size_t)nbk <= 150 * 1024,
                "%s: nb + kma too large for LDS", fn);
  const uint32_t mcount = (uint32_t)((float)n / (float)nthread);
  uint32_t nused = mcount * nthread;
  if (nused > n) nused = n;
  dim3 grid((nf + fpb - 1) / fpb, nq);
  const size_t shm = per_f * fpb + 4 * (size_t)nbk;
  if (ws) {
    unsigned char *w = static_cast<unsigned char *>(ws);
    uint32_t *maxbits = reinterpret_cast<uint32_t *>(w);
    float *lct = reinterpret_cast<float *>(w + bls_ws_part((size_t)nq * sizeof(uint32_t)));
    float *tut = simple ? nullptr
                        : reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(lct) +
                                                    bls_ws_part((size_t)n * nq * sizeof(float)));
    hipLaunchKernelGGL(bls_zero_kernel, dim3((nq + 255) / 256), dim3(256), 0, st, maxbits, nq);
    hipLaunchKernelGGL(bls_transpose_kernel,
                       dim3((nq + kBtTile - 1) / kBtTile, (n + kBtTile - 1) / kBtTile),
                       dim3(256), 0, st, lc, simple ? nullptr : tu, simple ? nullptr : offset,
                       lct, tut, maxbits, n, nq);
    if (simple)
      hipLaunchKernelGGL((bls_kernel<true, true>), grid, dim3(kBlsThreads), shm, st, blsp,
                         nullptr, nullptr, lct, tu, nullptr, nullptr, maxbits, n, nq, nf, nb,
                         kma, kmi, kkmi, fmin, df, nused, fpb);
    else
      hipLaunchKernelGGL((bls_kernel<false, true>), grid, dim3(kBlsThreads), shm, st, blsp,
                         phase1, phase2, lct, tut, offset, ntrue, maxbits, n, nq, nf, nb, kma,
                         kmi, kkmi, fmin, df, nused, fpb);
    return check_launch("bls_kernel");
  }
  if (simple)
    hipLaunchKernelGGL((bls_kernel<true, false>), grid, dim3(kBlsThreads), shm, st, blsp,
                       nullptr, nullptr, lc, tu, nullptr, nullptr, nullptr, n, nq, nf, nb, kma,
                       kmi, kkmi, fmin, df, nused, fpb);
  else
    hipLaunchKernelGGL((bls_kernel<false, false>), grid, dim3(kBlsThreads), shm, st, blsp,
                       phase1, phase2, lc, tu, offset, ntrue, nullptr, n, nq, nf, nb, kma, kmi,
                       kkmi, fmin, df, nused, fpb);
  return check_launch("bls_kernel");
}

}  // namespace gtrap

extern "C" int gtrap_geebls(float *blsp, uint32_t *phase1, uint32_t *phase2, const float *lc,
                            const float *tu, const float *offset, const uint32_t *ntrue,
                            uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma,
                            uint32_t kmi, uint32_t kkmi, float fmin, float df, uint32_t nthread,
                            void *stream) {
  return gtrap::bls_launch(false, blsp, phase1, phase2, lc, tu, offset, ntrue, n, nq, nf, nb,
                           kma, kmi, kkmi, fmin, df, nthread, nullptr, 0,
                           gtrap::as_stream(stream));
}

extern "C" int gtrap_geebls_simple(float *blsp, const float *lc, const float *tu, uint32_t n,
                                   uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma,
                                   uint32_t kmi, uint32_t kkmi, float fmin, float df,
                                   uint32_t nthread, void *stream) {
  return gtrap::bls_launch(true, blsp, nullptr, nullptr, lc, tu, nullptr, nullptr, n, nq, nf,
                           nb, kma, kmi, kkmi, fmin, df, nthread, nullptr, 0,
                           gtrap::as_stream(stream));
}

extern "C" int64_t gtrap_geebls_ws_bytes(uint32_t n, uint32_t nq, int32_t simple) {
  return (int64_t)gtrap::bls_ws_need(n, nq, simple != 0);
}

extern "C" int gtrap_geebls_ws(float *blsp, uint32_t *phase1, uint32_t *phase2, const float *lc,
                               const float *tu, const float *offset, const uint32_t *ntrue,
                               uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma,
                               uint32_t kmi, uint32_t kkmi, float fmin, float df,
                               uint32_t nthread, void *ws, int64_t ws_bytes, void *stream) {
  if (!ws) {
    gtrap::set_error("gtrap_geebls_ws: null workspace");
    return GTRAP_ENOMEM;
  }
  return gtrap::bls_launch(false, blsp, phase1, phase2, lc, tu, offset, ntrue, n, nq, nf, nb,
                           kma, kmi, kkmi, fmin, df, nthread, ws, ws_bytes,
                           gtrap::as_stream(stream));
}

extern "C" int gtrap_geebls_simple_ws(float *blsp, const float *lc, const float *tu, uint32_t n,
                                      uint32_t nq, uint32_t nf, uint32_t nb, uint32_t kma,
                                      uint32_t kmi, uint32_t kkmi, float fmin, float df,
                                      uint32_t nthread, void *ws, int64_t ws_bytes,
                                      void *stream) {
  if (!ws) {
    gtrap::set_error("gtrap_geebls_simple_ws: null workspace");
    return GTRAP_ENOMEM;
  }
  return gtrap::bls_launch(true, blsp, nullptr, nullptr, lc, tu, nullptr, nullptr, n, nq, nf,
                           nb, kma, kmi, kkmi, fmin, df, nthread, ws, ws_bytes,
                           gtrap::as_stream(stream));
}
