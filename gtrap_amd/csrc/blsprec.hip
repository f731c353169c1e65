// Arithmetic variants of geebls (include/geebls.h:1-167) for precision studies.
//
// Same box search as bls_kernel<false> (fold, bin, wrap the first kma bins,
// sequential running sums over widths, pow = s^2 / (kk (ntrue - kk)) with the
// kmi / kkmi thresholds, sqrt of the best), with three independent choices:
//
//   storage   F16: flux read as binary16 after (x - offset) * 2^k per curve
//             (the scale is a power of two, so the only error is the
//             binary16 rounding); else float32 lc - offset.
//   fold      FOLD64: f = fmin + df*jf and ph = frac(t*f) in float64 from a
//             float64 time array; else the reference's float32 fold of tu.
//   accum     0: float32 LDS atomics into the bins and a float32 running
//                sum (the reference's arithmetic; atomic order shows);
//             1: exact bin sums (64-bit fixed point) rounded to float32, a
//                float32 running sum (bls_kernel's arithmetic);
//             2: exact bin sums rounded to float32, a compensated (two-sum)
//                float32 running sum;
//             3: exact bin sums rounded to float64, float64 running sum and
//                statistic (with FOLD64 and float32 storage: the float64
//                result the others are measured against).
//
// Output SR in float64 (the float32 variants' values widened), and the
// best (start, end) bins with the lowest-index tie-break.
#include "common.h"
#include <hip/hip_fp16.h>
#include <math.h>

namespace gtrap {

constexpr int kBpThreads = 256;
constexpr int kBpSpt = 8;

__device__ __forceinline__ int bp_scale_exp(float maxabs, uint32_t n) {
  int e = -100;
  if (maxabs > 0.0f) frexpf(maxabs, &e);
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) <= (uint64_t)n) ++b;
  int S = 62 - e - b;
  return S > 200 ? 200 : S;
}

struct BpBest {
  double p;
  uint32_t i, j;
};

__device__ __forceinline__ bool bp_better(double p, uint32_t i, uint32_t j, const BpBest &b) {
  return (p > b.p) || (p == b.p && (i < b.i || (i == b.i && j < b.j)));
}

template <bool F16, int ACC, bool FOLD64>
__global__ __launch_bounds__(kBpThreads) void blsprec_kernel(
    double *__restrict__ sr, uint32_t *__restrict__ phase1, uint32_t *__restrict__ phase2,
    const float *__restrict__ lc, const __half *__restrict__ lch,
    const float *__restrict__ hinv, const float *__restrict__ tu,
    const double *__restrict__ t64, const float *__restrict__ offset,
    const uint32_t *__restrict__ ntrue, uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb,
    uint32_t kma, uint32_t kmi, uint32_t kkmi, float fmin, float df, double fmin64,
    double df64, uint32_t nused, int fpb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t nbk = nb + kma;
  // LDS: Y[fpb][nbk] int64 (float bins for ACC 0) | C[fpb][nbk] u32 | yd[nbk] double
  int64_t *Y = reinterpret_cast<int64_t *>(smem);
  uint32_t *Cn = reinterpret_cast<uint32_t *>(Y + (size_t)fpb * nbk);
  double *yd = reinterpret_cast<double *>(Cn + (size_t)fpb * nbk + ((size_t)fpb * nbk & 1));
  __shared__ double red_p[kBpThreads / kWave];
  __shared__ uint32_t red_i[kBpThreads / kWave], red_j[kBpThreads / kWave];
  __shared__ float red_max[kBpThreads / kWave];

  const uint32_t q = blockIdx.y;
  const uint32_t jf0 = blockIdx.x * (uint32_t)fpb;
  const int nfr = (int)min((uint32_t)fpb, nf - jf0);
  const int tid = threadIdx.x;
  const float off = offset[q];
  const float hs = F16 ? hinv[q] : 1.0f;
  const float rntrue = (float)ntrue[q];
  const double rntrue64 = (double)ntrue[q];
  const uint32_t kmi_e = kmi < 1 ? 1 : kmi;
  const uint32_t kkmi_e = kkmi < 5 ? 5 : kkmi;

  auto value = [&](uint32_t i) -> float {
    if (F16) return __half2float(lch[(size_t)i * nq + q]) * hs;  // exact power-of-two scale
    return lc[(size_t)i * nq + q] - off;
  };

  float mx = 0.0f;
  if (ACC != 0) {
    for (uint32_t i = tid; i < n; i += kBpThreads) mx = fmaxf(mx, fabsf(value(i)));
    for (int o = kWave / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((tid & (kWave - 1)) == 0) red_max[tid / kWave] = mx;
  }
  for (uint32_t e = tid; e < (uint32_t)fpb * nbk; e += kBpThreads) {
    Y[e] = 0;
    Cn[e] = 0;
  }
  __syncthreads();
  int S = 0;
  if (ACC != 0) {
    mx = red_max[0];
    for (int w = 1; w < kBpThreads / kWave; ++w) mx = fmaxf(mx, red_max[w]);
    S = bp_scale_exp(mx, n);
  }

  for (uint32_t base = 0; base < nused; base += kBpThreads * kBpSpt) {
    float tt[kBpSpt];
    double td[kBpSpt];
    float vv[kBpSpt];
    int64_t qv[kBpSpt];
#pragma unroll
    for (int s = 0; s < kBpSpt; ++s) {
      const uint32_t i = base + s * kBpThreads + tid;
      tt[s] = -1.0f;
      td[s] = -1.0;
      vv[s] = 0.0f;
      qv[s] = 0;
      if (i < nused) {
        tt[s] = tu[(size_t)i * nq + q];
        if (FOLD64) td[s] = t64[(size_t)i * nq + q];
        vv[s] = value(i);
        if (ACC != 0) qv[s] = llrint(ldexp((double)vv[s], S));
      }
    }
    for (int f = 0; f < nfr; ++f) {
      const uint32_t jf = jf0 + f;
      int64_t *Yf = Y + (size_t)f * nbk;
      uint32_t *Cf = Cn + (size_t)f * nbk;
#pragma unroll
      for (int s = 0; s < kBpSpt; ++s) {
        // validity (tu >= 0) from the stored float32 time in every variant
        const bool ok = (base + s * kBpThreads + tid < nused) && tt[s] >= 0.0f;
        if (!ok) continue;
        uint32_t j;
        if (FOLD64) {
          const double f0 = fmin64 + df64 * (double)jf;
          double ph = td[s] * f0;
          ph = ph - floor(ph);
          const double fj = (double)nb * ph;
          j = fj > 0.0 ? (uint32_t)fj : 0u;
        } else {
          const float f0 = fmaf(df, (float)jf, fmin);
          float ph = tt[s] * f0;
          ph = ph - (float)(int)ph;
          const float fj = (float)nb * ph;
          j = fj > 0.0f ? (uint32_t)fj : 0u;
        }
        if (j < nb) {
          if (ACC == 0)
            atomicAdd(reinterpret_cast<float *>(Yf) + j, vv[s]);
          else
            atomicAdd(reinterpret_cast<unsigned long long *>(Yf + j),
                      (unsigned long long)qv[s]);
          atomicAdd(Cf + j, 1u);
        }
      }
    }
  }
  __syncthreads();

  for (int f = 0; f < nfr; ++f) {
    int64_t *Yf = Y + (size_t)f * nbk;
    uint32_t *Cf = Cn + (size_t)f * nbk;
    for (uint32_t k = tid; k < kma; k += kBpThreads) {
      if (ACC == 0)
        reinterpret_cast<float *>(Yf)[nb + k] = reinterpret_cast<float *>(Yf)[k];
      else
        Yf[nb + k] = Yf[k];
      Cf[nb + k] = Cf[k];
    }
  }
  __syncthreads();

  for (int f = 0; f < nfr; ++f) {
    const int64_t *Yf = Y + (size_t)f * nbk;
    const uint32_t *Cf = Cn + (size_t)f * nbk;
    for (uint32_t e = tid; e < nbk; e += kBpThreads) {
      if (ACC == 0)
        yd[e] = (double)reinterpret_cast<const float *>(Yf)[e];
      else if (ACC == 3)
        yd[e] = ldexp((double)Yf[e], -S);
      else
        yd[e] = (double)(float)ldexp((double)Yf[e], -S);
    }
    __syncthreads();
    BpBest b{0.0, 0xFFFFFFFFu, 0xFFFFFFFFu};
    for (uint32_t i = tid; i < nb; i += kBpThreads) {
      float s = 0.0f, c = 0.0f;
      double sd = 0.0;
      uint32_t kk = 0;
      for (uint32_t k = 1; k <= kma; ++k) {
        const uint32_t j = i + k - 1;
        if (ACC == 3) {
          sd = sd + yd[j];
        } else if (ACC == 2) {
          const float y = (float)yd[j];
          const float t = s + y;
          const float z = t - s;
          c = c + ((s - (t - z)) + (y - z));
          s = t;
        } else {
          s = s + (float)yd[j];
        }
        kk = kk + Cf[j];
        if (k > kmi_e && kk > kkmi_e) {
          double pw;
          if (ACC == 3) {
            const double rn1 = (double)kk;
            pw = (sd * sd) / (rn1 * (rntrue64 - rn1));
          } else {
            const float sv = ACC == 2 ? s + c : s;
            const float rn1 = (float)kk;
            pw = (double)((sv * sv) / (rn1 * (rntrue - rn1)));
          }
          if (pw > b.p) { b.p = pw; b.i = i; b.j = j; }
        }
      }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const double p2 = __shfl_xor(b.p, o);
      const uint32_t i2 = __shfl_xor(b.i, o), j2 = __shfl_xor(b.j, o);
      if (bp_better(p2, i2, j2, b)) { b.p = p2; b.i = i2; b.j = j2; }
    }
    if ((tid & (kWave - 1)) == 0) {
      red_p[tid / kWave] = b.p;
      red_i[tid / kWave] = b.i;
      red_j[tid / kWave] = b.j;
    }
    __syncthreads();
    if (tid == 0) {
      BpBest r{red_p[0], red_i[0], red_j[0]};
      for (int w = 1; w < kBpThreads / kWave; ++w)
        if (bp_better(red_p[w], red_i[w], red_j[w], r)) r = BpBest{red_p[w], red_i[w], red_j[w]};
      const size_t o = (size_t)(jf0 + f) + (size_t)nf * q;
      sr[o] = ACC == 3 ? sqrt(r.p) : (double)sqrtf((float)r.p);
      if (phase1) phase1[o] = r.i;
      if (phase2) phase2[o] = r.j;
    }
    __syncthreads();
  }
}

template <bool F16, int ACC, bool FOLD64>
static void bp_launch(dim3 grid, size_t shm, hipStream_t st, double *sr, uint32_t *p1,
                      uint32_t *p2, const float *lc, const __half *lch, const float *hinv,
                      const float *tu, const double *t64, const float *offset,
                      const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf, uint32_t nb,
                      uint32_t kma, uint32_t kmi, uint32_t kkmi, float fmin, float df,
                      double fmin64, double df64, uint32_t nused, int fpb) {
  hipLaunchKernelGGL((blsprec_kernel<F16, ACC, FOLD64>), grid, dim3(kBpThreads), shm, st, sr,
                     p1, p2, lc, lch, hinv, tu, t64, offset, ntrue, n, nq, nf, nb, kma, kmi,
                     kkmi, fmin, df, fmin64, df64, nused, fpb);
}

}  // namespace gtrap

extern "C" int gtrap_geebls_prec(int variant, double *sr, uint32_t *phase1, uint32_t *phase2,
                                 const float *lc, const uint16_t *lch, const float *hinv,
                                 const float *tu, const double *t64, const float *offset,
                                 const uint32_t *ntrue, uint32_t n, uint32_t nq, uint32_t nf,
                                 uint32_t nb, uint32_t kma, uint32_t kmi, uint32_t kkmi,
                                 double fmin, double df, uint32_t nthread, void *stream) {
  using namespace gtrap;
  const bool f16 = variant & 1;
  const int acc = (variant >> 1) & 3;
  const bool fold64 = (variant >> 3) & 1;
  GTRAP_REQUIRE(variant >= 0 && variant < 16, "gtrap_geebls_prec: variant must be 0..15");
  GTRAP_REQUIRE(sr && tu && offset && ntrue, "gtrap_geebls_prec: null pointer");
  GTRAP_REQUIRE(f16 ? (lch && hinv) : lc != nullptr, "gtrap_geebls_prec: flux input missing");
  GTRAP_REQUIRE(!fold64 || t64, "gtrap_geebls_prec: float64 time required for the float64 fold");
  GTRAP_REQUIRE(nthread >= 1 && nthread <= 1024, "gtrap_geebls_prec: nthread must be 1..1024");
  GTRAP_REQUIRE(nb >= 1 && (nb & (nb - 1)) == 0 && nb % nthread == 0,
                "gtrap_geebls_prec: nb must be a power of two and a multiple of nthread");
  GTRAP_REQUIRE(kma >= 1 && kma <= nthread && kma <= nb, "gtrap_geebls_prec: bad kma");
  GTRAP_REQUIRE(n >= 1 && nq >= 1 && nf >= 1 && nq <= 65535, "gtrap_geebls_prec: bad sizes");
  const uint32_t nbk = nb + kma;
  const size_t per_f = (size_t)nbk * 12;
  const size_t fixed = 8 * (size_t)nbk + 8;
  int fpb = (int)((64 * 1024 - fixed) / per_f);
  if (fpb > 8) fpb = 8;
  GTRAP_REQUIRE(fpb >= 1, "gtrap_geebls_prec: nb + kma too large for LDS");
  const uint32_t mcount = (uint32_t)((float)n / (float)nthread);
  uint32_t nused = mcount * nthread;
  if (nused > n) nused = n;
  dim3 grid((nf + fpb - 1) / fpb, nq);
  const size_t shm = per_f * fpb + fixed;
  hipStream_t st = as_stream(stream);
  const __half *h = reinterpret_cast<const __half *>(lch);
  const float f32min = (float)fmin, f32df = (float)df;
#define GTRAP_BP(F, A, D)                                                                     \
  if (f16 == F && acc == A && fold64 == D)                                                    \
    bp_launch<F, A, D>(grid, shm, st, sr, phase1, phase2, lc, h, hinv, tu, t64, offset,       \
                       ntrue, n, nq, nf, nb, kma, kmi, kkmi, f32min, f32df, fmin, df, nused,  \
                       fpb);
  GTRAP_BP(false, 0, false) GTRAP_BP(false, 1, false) GTRAP_BP(false, 2, false)
  GTRAP_BP(false, 3, false) GTRAP_BP(true, 0, false) GTRAP_BP(true, 1, false)
  GTRAP_BP(true, 2, false) GTRAP_BP(true, 3, false) GTRAP_BP(false, 0, true)
  GTRAP_BP(false, 1, true) GTRAP_BP(false, 2, true) GTRAP_BP(false, 3, true)
  GTRAP_BP(true, 0, true) GTRAP_BP(true, 1, true) GTRAP_BP(true, 2, true)
  GTRAP_BP(true, 3, true)
#undef GTRAP_BP
  return check_launch("blsprec_kernel");
}
